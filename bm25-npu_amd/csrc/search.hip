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
//   k_score  - persistent workgroups walk (query, doc-block) tasks. Each task streams the
//              query terms' posting slices for its block (from a block directory, or a
//              two-round 64-ary wave search for short lists) with 8-byte {doc, impact}
//              loads and adds impacts into LDS FP32 accumulators, term by term in query
//              order with a barrier between terms, so every doc's score is the same
//              sequential FP32 sum as the CPU reference (cpu_index_search.h:176-189).
//              Then it keeps the block's exact top-K among docs scoring at least a
//              per-query running lower bound of the global K-th score, and appends them
//              as 64-bit keys (score bits << 32 | ~doc) to the query's candidate pool.
//   k_final  - one workgroup per query selects the exact top-K of its pool by a 64-step
//              bisection on the composite key (unique per doc, so ties need no extra
//              pass), bitonic-sorts it (score desc, doc asc: topk_heap.h:11-16) and
//              writes it to pinned host memory. Fewer than K positive docs are padded
//              with the lowest-ID zero-score docs, as the reference does.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace bm25mi {

__device__ __forceinline__ uint32_t WaveSum(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ uint32_t WaveMax(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = max(v, static_cast<uint32_t>(__shfl_xor(v, o)));
    return v;
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
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o)
            x += y;
    }
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

// lower_bound of a and of b (a <= b) among doc IDs post[lo, hi), hi - lo < 4096.
// Executed by one full wave: round 1 samples every ceil(len/64)-th posting, round 2
// resolves the <= 63-posting window each sample count points at.
__device__ __forceinline__ void WaveBounds(const uint2 *post, uint32_t lo, uint32_t hi, int32_t a, int32_t b,
                                           uint32_t &ra, uint32_t &rb)
{
    const uint32_t lane = __lane_id();
    const uint32_t len = hi - lo;
    const uint32_t step = (len + 63) / 64;
    const uint32_t p = lo + lane * step;
    const bool valid = lane * step < len;
    const int32_t v = valid ? static_cast<int32_t>(post[p].x) : 0;
    const uint32_t ca = __popcll(__ballot(valid && v < a));
    const uint32_t cb = __popcll(__ballot(valid && v < b));
    uint32_t r0a = lo, r1a = lo, r0b = lo, r1b = lo;
    if (ca > 0) {
        r0a = lo + (ca - 1) * step + 1;
        r1a = min(lo + ca * step, hi);
    }
    if (cb > 0) {
        r0b = lo + (cb - 1) * step + 1;
        r1b = min(lo + cb * step, hi);
    }
    const bool va = r0a + lane < r1a, vb = r0b + lane < r1b;
    const int32_t xa = va ? static_cast<int32_t>(post[r0a + lane].x) : 0;
    const int32_t xb = vb ? static_cast<int32_t>(post[r0b + lane].x) : 0;
    ra = r0a + __popcll(__ballot(va && xa < a));
    rb = r0b + __popcll(__ballot(vb && xb < b));
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

// Lower bound of a query's K-th score from its kept-candidate histogram, computed by one
// wave: the edge of the highest bucket b with at least k counted scores at or above it
// (those k candidates score at least that edge). 1 when fewer than k are counted.
__device__ __forceinline__ uint32_t WaveHistBound(const uint32_t *hist, uint32_t k)
{
    constexpr int kPer = (kHistBins + 63) / 64;
    const int lane = static_cast<int>(__lane_id());
    uint32_t h[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int b = lane * kPer + i;
        h[i] = b < kHistBins ? __atomic_load_n(hist + b, __ATOMIC_RELAXED) : 0u;
        sum += h[i];
    }
    // above = counts in buckets of higher lanes (suffix over lanes).
    uint32_t x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_down(x, o);
        if (lane + o < 64)
            x += y;
    }
    uint32_t above = x - sum;
    uint32_t best = 0; // bucket + 1, or 0
#pragma unroll
    for (int i = kPer - 1; i >= 0; --i) {
        above += h[i];
        if (best == 0 && above >= k)
            best = static_cast<uint32_t>(lane * kPer + i) + 1;
    }
    const uint32_t top = WaveMax(best);
    return top ? HistEdge(top - 1) : 1u;
}

// Planning kernel, one wave per (query, block) task. Finds the posting slice of every
// query word inside the block (block directory for long lists, a two-round 64-ary search
// for short ones) and writes the task's posting segments:
// * plain queries: one segment per word (its block slice);
// * threshold queries: an upper bound of every 256-doc granule's scores (per-word granule
//   maxima summed in word order; short words count their largest impact), then per word
//   only the runs of granules whose bound reaches the threshold. Docs in other granules
//   can only collect partial sums below the threshold, so they are never kept. A task with
//   no such granule gets no segment; one needing more than kMaxSegs keeps whole slices.
__global__ __launch_bounds__(256) void KPlan(IndexView ix, BatchView bt, WorkView wk, int32_t blocks, int32_t blockDocs,
                                             int32_t granulesPerBlock)
{
    const int lane = threadIdx.x & 63;
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (blockDim.x / 64);
    const int32_t nq = bt.header[0];
    const int64_t total = static_cast<int64_t>(nq) * blocks;
    for (int64_t task = static_cast<int64_t>(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6); task < total;
         task += waves) {
        const int q = static_cast<int>(task % nq);
        const int blk = static_cast<int>(task / nq);
        const int32_t base = blk * blockDocs;
        const int32_t end = min(base + blockDocs, ix.ndocs);
        const int64_t w0 = bt.qoff[q];
        const int n = static_cast<int>(bt.qoff[q + 1] - w0);
        const float thr = bt.thr[q];
        const int64_t slotOut = static_cast<int64_t>(q) * blocks + blk;
        uint2 *segs = wk.segs + slotOut * kMaxSegs;
        uint4 info = make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
        if (lane < n)
            info = bt.words[w0 + lane];
        uint32_t lo = 0, hi = 0;
        const int32_t sl = static_cast<int32_t>(info.z);
        const int64_t g0 = static_cast<int64_t>(blk) * granulesPerBlock;
        const int64_t g1 = min(g0 + granulesPerBlock, ix.granules);
        if (info.y > 0 && sl >= 0) {
            const uint32_t *d = ix.dir + static_cast<int64_t>(sl) * (ix.granules + 1);
            lo = info.x + d[g0];
            hi = info.x + d[g1];
        }
        // Short lists: the whole wave searches each one in turn.
        unsigned long long shortMask = __ballot(lane < n && info.y > 0 && sl < 0);
        while (shortMask) {
            const int j = __ffsll(static_cast<long long>(shortMask)) - 1;
            shortMask &= shortMask - 1;
            const uint32_t first = __shfl(info.x, j), count = __shfl(info.y, j);
            uint32_t a, b;
            WaveBounds(ix.post, first, first + count, base, end, a, b);
            if (lane == j) {
                lo = a;
                hi = b;
            }
        }
        if (thr <= 0.0F) {
            if (lane < n)
                segs[lane] = make_uint2(lo, ((hi - lo) << 8) | static_cast<uint32_t>(lane));
            if (lane == 0)
                wk.segCount[slotOut] = static_cast<uint32_t>(n);
            continue;
        }
        // Granule bounds (lane g: granule g0 + g), summed in word order.
        const int gpb = static_cast<int>(g1 - g0);
        float ub = 0.0F;
        for (int j = 0; j < n; ++j) {
            const uint32_t jlo = __shfl(lo, j), jhi = __shfl(hi, j);
            const int32_t jsl = static_cast<int32_t>(__shfl(info.z, j));
            const float jmax = __uint_as_float(__shfl(info.w, j));
            if (jhi > jlo && lane < gpb)
                ub += jsl >= 0 ? ix.bmax[static_cast<int64_t>(jsl) * ix.granules + g0 + lane] : jmax;
        }
        const unsigned long long active = __ballot(lane < gpb && ub >= thr);
        uint32_t nseg = 0;
        if (active) {
            // Runs of active granules: [start bit, first inactive bit after it).
            bool overflow = false;
            for (int j = 0; j < n && !overflow; ++j) {
                const uint32_t jlo = __shfl(lo, j), jhi = __shfl(hi, j);
                if (jhi <= jlo)
                    continue;
                const int32_t jsl = static_cast<int32_t>(__shfl(info.z, j));
                if (jsl < 0) {
                    // Short word: its whole block slice.
                    if (nseg >= static_cast<uint32_t>(kMaxSegs)) {
                        overflow = true;
                        break;
                    }
                    if (lane == 0)
                        segs[nseg] = make_uint2(jlo, ((jhi - jlo) << 8) | static_cast<uint32_t>(j));
                    ++nseg;
                    continue;
                }
                const uint32_t *d = ix.dir + static_cast<int64_t>(jsl) * (ix.granules + 1);
                const uint32_t first = __shfl(info.x, j);
                const uint32_t dv = lane <= gpb ? d[g0 + lane] : 0u; // lane g: postings before granule g
                unsigned long long runs = active & ~(active << 1);      // run starts
                while (runs) {
                    const int ga = __ffsll(static_cast<long long>(runs)) - 1;
                    runs &= runs - 1;
                    const unsigned long long after = ~active & (~0ULL << ga);
                    const int gb = after ? min(__ffsll(static_cast<long long>(after)) - 1, gpb) : gpb;
                    const uint32_t s0 = first + __shfl(dv, ga), s1 = first + __shfl(dv, gb);
                    if (s1 > s0) {
                        if (nseg >= static_cast<uint32_t>(kMaxSegs)) {
                            overflow = true;
                            break;
                        }
                        if (lane == 0)
                            segs[nseg] = make_uint2(s0, ((s1 - s0) << 8) | static_cast<uint32_t>(j));
                        ++nseg;
                    }
                }
            }
            if (overflow) {
                // Too fragmented: whole block slices instead.
                if (lane < n)
                    segs[lane] = make_uint2(lo, ((hi - lo) << 8) | static_cast<uint32_t>(lane));
                nseg = static_cast<uint32_t>(n);
            }
        }
        if (lane == 0)
            wk.segCount[slotOut] = nseg;
    }
}

// The scoring kernel; persistent workgroups walk (query, block) tasks, every
// kPilotStride-th block first so that the first round samples the whole corpus. Each task
// keeps its block's top-K among docs scoring at least the query's lower bound tau and
// counts them in the query's score histogram. tau is the larger of the best block K-th
// score seen so far and the histogram bound (the K-th best kept score, to 1/16 octave),
// so after the first round few docs qualify and most tasks keep nothing.
template <int B>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void KScore(IndexView ix, BatchView bt,
                                                                                          WorkView wk, int32_t blocks)
{
    constexpr int kRows = B / (4 * kThreads); // float4 rows per thread
    constexpr int U = kLoadsPerThread;        // 256-posting columns per round
    __shared__ __align__(16) float acc[B];
    __shared__ uint32_t sLo[kMaxSegs], sWord[kMaxSegs], sPre[kMaxSegs + 1], sCol[kMaxSegs + 1];
    __shared__ uint2 sRound[U];
    __shared__ uint32_t sRed[kWaves];
    __shared__ uint32_t sShared[4];
    __shared__ unsigned long long sTop[kThreads]; // small exact selections (also per-thread maxima)

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t nq = bt.header[0];
    const int32_t firstBlocks = (blocks + kPilotStride - 1) / kPilotStride;
    const int64_t firstTasks = static_cast<int64_t>(nq) * firstBlocks;
    const int64_t total = static_cast<int64_t>(nq) * blocks;
    // Task order: blocks 0, s, 2s, ... for every query, then the other blocks.
    auto blockOf = [&](int64_t t) {
        if (t < firstTasks)
            return static_cast<int>(t / nq) * kPilotStride;
        const int64_t j = (t - firstTasks) / nq;
        return static_cast<int>(kPilotStride * (j / (kPilotStride - 1)) + j % (kPilotStride - 1) + 1);
    };
    auto queryOf = [&](int64_t t) { return static_cast<int>((t < firstTasks ? t : t - firstTasks) % nq); };
    // This lane's posting segment of a task (lane s: segment s) and the task's segment count.
    auto segOf = [&](int64_t t, uint32_t &count) {
        uint2 v = make_uint2(0u, 0u);
        count = 0;
        if (t < total) {
            const int64_t at = static_cast<int64_t>(queryOf(t)) * blocks + blockOf(t);
            count = wk.segCount[at];
            v = wk.segs[at * kMaxSegs + lane];
        }
        return v;
    };
    unsigned long long scanned = 0, scored = 0, skipped = 0, emitted = 0, overflow = 0, emitting = 0, exactTasks = 0;
#ifdef BM25MI_PHASES
    // Cycle counters of thread 0 per phase: setup, postings, selection, rest.
    unsigned long long phase[4] = {0, 0, 0, 0};
    unsigned long long stamp = clock64();
#define BM25MI_PHASE(i)                                                                                                \
    do {                                                                                                               \
        if (tid == 0) {                                                                                                \
            const unsigned long long now_ = clock64();                                                                 \
            phase[i] += now_ - stamp;                                                                                  \
            stamp = now_;                                                                                              \
        }                                                                                                              \
    } while (0)
#else
#define BM25MI_PHASE(i)                                                                                                \
    do {                                                                                                               \
    } while (0)
#endif

    for (int64_t task = blockIdx.x; task < total; task += gridDim.x) {
        const int q = queryOf(task);
        const int blk = blockOf(task);
        const float thr = bt.thr[q];
        const int32_t base = blk * B;
        const int32_t k = bt.k[q];
        __syncthreads(); // previous task is done with LDS

        if (wave == 0) {
            // This task's posting segments (lane s: segment s; loaded during the previous
            // task), then prefixes of segment lengths and of 256-posting column counts.
            uint32_t nseg;
            const uint2 sg = segOf(task, nseg);
            const uint32_t len = lane < static_cast<int>(nseg) ? (sg.y >> 8) : 0u;
            uint32_t x = len, c = (len + kThreads - 1) / kThreads;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o), z = __shfl_up(c, o);
                if (lane >= o) {
                    x += y;
                    c += z;
                }
            }
            if (lane < static_cast<int>(nseg)) {
                sLo[lane] = sg.x;
                sWord[lane] = sg.y & 0xFFu;
                sPre[lane + 1] = x;
                sCol[lane + 1] = c;
            }
            const uint32_t hb = WaveHistBound(wk.hist + static_cast<int64_t>(q) * kHistBins, static_cast<uint32_t>(k));
            if (lane == 0) {
                sPre[0] = 0;
                sCol[0] = 0;
                sShared[0] = max(max(__atomic_load_n(&wk.tau[q], __ATOMIC_RELAXED), hb), 1u);
                sShared[2] = (thr > 0.0F && nseg == 0) ? 0u : 1u;
                sShared[3] = nseg;
            }
        }
        float4 *acc4 = reinterpret_cast<float4 *>(acc);
#pragma unroll
        for (int r = 0; r < kRows; ++r)
            acc4[r * kThreads + tid] = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
        __syncthreads();
        BM25MI_PHASE(0);
        if (sShared[2] == 0) {
            ++skipped; // every granule's score upper bound is below the threshold
            continue;
        }
        ++scored;
        uint32_t tau = sShared[0];
        if (thr > 0.0F)
            tau = max(tau, __float_as_uint(thr));

        // Term-ordered FP32 accumulation. Each word's slice is cut into 256-posting
        // columns; a round issues up to U columns' loads (one 8-byte load per thread per
        // column) before any add, then adds them column by column with a barrier where the
        // word changes. Within one word each doc appears once, so lanes never collide, and
        // every doc's score is summed in query-word order.
        const uint32_t nseg = sShared[3];
        const uint32_t columns = sCol[nseg];
        if (tid == 0)
            scanned += sPre[nseg];
        for (uint32_t r0 = 0; r0 < columns; r0 += U) {
            if (wave == 0 && lane < U) {
                const uint32_t col = r0 + lane;
                uint2 entry = make_uint2(0u, 0xFFFF0000u);
                if (col < columns) {
                    int sg = 0;
                    while (sCol[sg + 1] <= col)
                        ++sg;
                    const uint32_t i = col - sCol[sg];
                    const uint32_t len = sPre[sg + 1] - sPre[sg];
                    entry = make_uint2(sLo[sg] + i * kThreads, min(static_cast<uint32_t>(kThreads), len - i * kThreads) |
                                                                   (sWord[sg] << 16));
                }
                sRound[lane] = entry;
            }
            __syncthreads();
            uint2 e[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint2 c = sRound[u];
                if (static_cast<uint32_t>(tid) < (c.y & 0xFFFFu))
                    e[u] = ix.post[c.x + tid];
            }
            // Unused columns carry word 0xFFFF and count 0: no barrier, no add. The barrier
            // condition reads LDS values shared by the whole workgroup.
            uint32_t word = sRound[0].y >> 16;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t meta = sRound[u].y;
                const uint32_t w = meta >> 16;
                if (w != 0xFFFFu && w != word) {
                    __syncthreads();
                    word = w;
                }
                if (static_cast<uint32_t>(tid) < (meta & 0xFFFFu))
                    acc[static_cast<int32_t>(e[u].x) - base] += __uint_as_float(e[u].y);
            }
            __syncthreads();
        }
        BM25MI_PHASE(1);

        // Block selection. Scores are re-read from LDS in each pass (8 x 16-byte reads per
        // thread) so they do not occupy registers across the passes.
        const uint4 *bits4 = reinterpret_cast<const uint4 *>(acc);
        auto countIf = [&](auto pred) {
            uint32_t c = 0;
#pragma unroll 2
            for (int r = 0; r < kRows; ++r) {
                const uint4 v = bits4[r * kThreads + tid];
                const uint32_t e = r * 4 * kThreads + tid * 4;
                c += (pred(v.x, e) ? 1u : 0u) + (pred(v.y, e + 1) ? 1u : 0u) + (pred(v.z, e + 2) ? 1u : 0u) +
                     (pred(v.w, e + 3) ? 1u : 0u);
            }
            return c;
        };
        auto countGe = [&](uint32_t x) {
            return BlockSum<kThreads>(countIf([x](uint32_t b, uint32_t) { return b >= x; }), sRed);
        };
        auto keyOf = [base](uint32_t b, uint32_t e) {
            return (static_cast<unsigned long long>(b) << 32) |
                   static_cast<unsigned long long>(0xFFFFFFFFu - (static_cast<uint32_t>(base) + e));
        };
        unsigned long long *pool = wk.pool + bt.poolOff[q];
        uint32_t *hist = wk.hist + static_cast<int64_t>(q) * kHistBins;
        const int64_t cap = bt.poolOff[q + 1] - bt.poolOff[q];
        const uint32_t kk = static_cast<uint32_t>(k);
        uint32_t floor = tau;
        uint32_t mine = 0, mx = 0;
#pragma unroll 2
        for (int r = 0; r < kRows; ++r) {
            const uint4 v = bits4[r * kThreads + tid];
            mine += (v.x >= floor ? 1u : 0u) + (v.y >= floor ? 1u : 0u) + (v.z >= floor ? 1u : 0u) +
                    (v.w >= floor ? 1u : 0u);
            mx = max(mx, max(max(v.x, v.y), max(v.z, v.w)));
        }
        uint32_t candidates = BlockSum<kThreads>(mine, sRed);
        if (candidates > 0) {
            uint32_t cut = floor, tieIdx = B; // keep bits > cut, or bits == cut at index <= tieIdx
            uint32_t kth = 0;                 // the block's K-th score bits when it has more than K
            bool sorted = false;
            if (candidates > kk && kk <= static_cast<uint32_t>(kThreads)) {
                if (candidates > static_cast<uint32_t>(kThreads)) {
                    // Raise the floor to the kk-th largest per-thread maximum: kk docs reach
                    // it, so the block's kk-th score is at least that.
                    sTop[tid] = mx;
                    __syncthreads();
                    BitonicDesc<kThreads>(sTop, kThreads);
                    floor = max(floor, static_cast<uint32_t>(sTop[kk - 1]));
                    __syncthreads();
                    mine = countIf([floor](uint32_t b, uint32_t) { return b >= floor; });
                    candidates = BlockSum<kThreads>(mine, sRed);
                }
                if (candidates <= static_cast<uint32_t>(kThreads)) {
                    // Few enough to sort: compact their composite keys and sort them; the
                    // first kk are the block's top-kk.
                    uint32_t count = 0;
                    uint32_t at = BlockExclusive<kThreads>(mine, sRed, &count);
#pragma unroll 1
                    for (int r = 0; r < kRows; ++r) {
                        const uint4 v = bits4[r * kThreads + tid];
                        const uint32_t e = r * 4 * kThreads + tid * 4;
                        if (v.x >= floor)
                            sTop[at++] = keyOf(v.x, e);
                        if (v.y >= floor)
                            sTop[at++] = keyOf(v.y, e + 1);
                        if (v.z >= floor)
                            sTop[at++] = keyOf(v.z, e + 2);
                        if (v.w >= floor)
                            sTop[at++] = keyOf(v.w, e + 3);
                    }
                    __syncthreads();
                    if (static_cast<uint32_t>(tid) >= count)
                        sTop[tid] = 0ULL;
                    __syncthreads();
                    BitonicDesc<kThreads>(sTop, kThreads);
                    kth = static_cast<uint32_t>(sTop[kk - 1] >> 32);
                    sorted = true;
                }
            }
            if (!sorted && candidates > kk) {
                // General exact cut: kk-th largest score bits by bisection,
                // count(>= lo) >= kk > count(>= hi), then ties by lowest doc ID.
                mx = BlockMax<kThreads>(mx, sRed);
                uint32_t lo = floor, hi = mx + 1;
                while (hi - lo > 1) {
                    const uint32_t mid = lo + (hi - lo) / 2;
                    if (countGe(mid) >= kk)
                        lo = mid;
                    else
                        hi = mid;
                }
                cut = lo;
                kth = cut;
                const uint32_t need = kk - countGe(lo + 1);
                const uint32_t c0 = cut;
                const uint32_t ties = BlockSum<kThreads>(countIf([c0](uint32_t b, uint32_t) { return b == c0; }), sRed);
                if (ties > need) {
                    // Smallest local index x with #{ties at index <= x} >= need.
                    int32_t a = -1, z = B - 1; // count(<= a) < need <= count(<= z)
                    while (z - a > 1) {
                        const int32_t mid = a + (z - a) / 2;
                        const uint32_t c = countIf([c0, mid](uint32_t b, uint32_t e) {
                            return b == c0 && static_cast<int32_t>(e) <= mid;
                        });
                        if (BlockSum<kThreads>(c, sRed) >= need)
                            z = mid;
                        else
                            a = mid;
                    }
                    tieIdx = static_cast<uint32_t>(z);
                }
            }
            // Append the kept docs and count them in the histogram.
            uint32_t kept;
            if (sorted) {
                kept = kk;
                if (tid == 0)
                    sShared[1] = atomicAdd(&wk.count[q], kk);
                __syncthreads();
                if (static_cast<uint32_t>(tid) < kk) {
                    const uint32_t slot = sShared[1] + tid;
                    if (static_cast<int64_t>(slot) < cap)
                        pool[slot] = sTop[tid];
                    else
                        ++overflow;
                    atomicAdd(hist + HistBucket(static_cast<uint32_t>(sTop[tid] >> 32)), 1u);
                }
            } else {
                auto keepIt = [cut, tieIdx](uint32_t b, uint32_t e) { return b > cut || (b == cut && e <= tieIdx); };
                mine = countIf(keepIt);
                uint32_t before = BlockExclusive<kThreads>(mine, sRed, &kept);
                if (tid == 0)
                    sShared[1] = atomicAdd(&wk.count[q], kept);
                __syncthreads();
                if (mine > 0) {
                    uint32_t slot = sShared[1] + before;
#pragma unroll 1
                    for (int r = 0; r < kRows; ++r) {
                        const uint4 v = bits4[r * kThreads + tid];
                        const uint32_t e = r * 4 * kThreads + tid * 4;
                        const uint32_t vals[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                        for (int cc = 0; cc < 4; ++cc)
                            if (keepIt(vals[cc], e + cc)) {
                                if (static_cast<int64_t>(slot) < cap)
                                    pool[slot] = keyOf(vals[cc], e + cc);
                                else
                                    ++overflow;
                                ++slot;
                                atomicAdd(hist + HistBucket(vals[cc]), 1u);
                            }
                    }
                }
            }
            if (kth) {
                ++exactTasks;
                if (tid == 0)
                    atomicMax(&wk.tau[q], kth); // the block's K-th score bounds the query's
            }
            if (tid == 0 && kept) {
                emitted += kept;
                ++emitting;
            }
        }
        BM25MI_PHASE(2);
    }
#ifdef BM25MI_PHASES
    if (tid == 0)
        for (int i = 0; i < 3; ++i)
            atomicAdd(&wk.stats[kStatPhase + i], phase[i]);
#endif
#undef BM25MI_PHASE
    if (tid == 0) {
        atomicAdd(&wk.stats[kStatPostings], scanned);
        atomicAdd(&wk.stats[kStatScored], scored);
        atomicAdd(&wk.stats[kStatSkipped], skipped);
        atomicAdd(&wk.stats[kStatCandidates], emitted);
        atomicAdd(&wk.stats[kStatEmitting], emitting);
        atomicAdd(&wk.stats[kStatExact], exactTasks);
    }
    if (overflow)
        atomicAdd(&wk.stats[kStatOverflow], overflow);
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
    if (i < kStats)
        wk.stats[i] = 0;
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

// Per query: the exact top-K of the candidate pool. Entries scoring below the query's
// final lower bound tau cannot be in the top-K and are dropped; the rest are sorted in
// LDS (score desc, doc asc) when they fit, else reduced by a 64-step bisection first.
// Fewer than K positive docs are padded with the lowest-ID zero-score docs.
__global__ __launch_bounds__(kFinalThreads) void KFinal(BatchView bt, WorkView wk, unsigned long long *hostStats)
{
    __shared__ unsigned long long sKeys[kSortCap];
    __shared__ unsigned long long sIds[1024];
    __shared__ uint32_t sRed[kFinalThreads / 64];
    __shared__ uint32_t sN;
    const int q = blockIdx.x;
    const int32_t nq = bt.header[0];
    if (q >= nq)
        return;
    const int tid = threadIdx.x;
    if (q == 0 && tid < kStats) {
        // The scoring kernel has finished: publish its counters and clear them.
        hostStats[tid] = wk.stats[tid];
        wk.stats[tid] = 0;
    }
    const uint32_t k = static_cast<uint32_t>(bt.k[q]);
    const int64_t cap = bt.poolOff[q + 1] - bt.poolOff[q];
    const uint32_t p = static_cast<uint32_t>(min(static_cast<int64_t>(wk.count[q]), cap));
    uint32_t *hist = wk.hist + static_cast<int64_t>(q) * kHistBins;
    if (tid < 64)
        sRed[0] = 0; // placeholder write order: the bound below is computed by wave 0
    __syncthreads();
    if (tid < 64) {
        const uint32_t hb = WaveHistBound(hist, k);
        if (tid == 0)
            sRed[0] = max(max(wk.tau[q], hb), 1u);
    }
    __syncthreads();
    const uint32_t tau = sRed[0];
    const unsigned long long *pool = wk.pool + bt.poolOff[q];
    const bool pruned = bt.thr[q] > 0.0F;
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
    for (uint32_t i0 = 0; i0 < p; i0 += kFinalThreads) {
        const uint32_t i = i0 + tid;
        const unsigned long long key = i < p ? pool[i] : 0ULL;
        const uint32_t at = WaveAppend(i < p && static_cast<uint32_t>(key >> 32) >= tau, &sN);
        if (at < static_cast<uint32_t>(kSortCap))
            sKeys[at] = key;
    }
    __syncthreads();
    const uint32_t c = sN;
    if (pruned && c < k) {
        if (tid == 0)
            wk.status[q] = 1; // fewer than K survivors: the host reruns this query on the plain path
        return;
    }
    const uint32_t m = c < k ? c : k;
    if (c <= static_cast<uint32_t>(kSortCap)) {
        int n2 = 1;
        while (n2 < static_cast<int>(c))
            n2 <<= 1;
        for (int i = static_cast<int>(c) + tid; i < n2; i += kFinalThreads)
            sKeys[i] = 0ULL;
        __syncthreads();
        BitonicDesc<kFinalThreads>(sKeys, n2);
    } else {
        // Too many to sort: the K-th key over the whole pool, then the K keys at or above it.
        const unsigned long long kappa = KthFromGlobal(pool, p, k, sRed);
        if (tid == 0)
            sN = 0;
        __syncthreads();
        for (uint32_t i = tid; i < p; i += kFinalThreads)
            if (pool[i] >= kappa) {
                const uint32_t at = atomicAdd(&sN, 1u);
                if (at < k)
                    sIds[at] = pool[i];
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
    const int grid = static_cast<int>((max(static_cast<int64_t>(maxQ) * kHistBins, static_cast<int64_t>(kStats)) +
                                       threads - 1) / threads);
    hipLaunchKernelGGL(KReset, dim3(grid), dim3(threads), 0, stream, wk, maxQ);
}

void LaunchPlan(const IndexView &ix, const BatchView &bt, const WorkView &wk, int32_t blocks, int32_t blockDocs,
                int32_t granulesPerBlock, int32_t grid, hipStream_t stream)
{
    hipLaunchKernelGGL(KPlan, dim3(grid), dim3(256), 0, stream, ix, bt, wk, blocks, blockDocs, granulesPerBlock);
}

bool LaunchScore(int blockDocs, const IndexView &ix, const BatchView &bt, const WorkView &wk, int32_t blocks,
                 int32_t grid, hipStream_t stream)
{
    switch (blockDocs) {
    case 2048:
        hipLaunchKernelGGL(KScore<2048>, dim3(grid), dim3(kThreads), 0, stream, ix, bt, wk, blocks);
        return true;
    case 4096:
        hipLaunchKernelGGL(KScore<4096>, dim3(grid), dim3(kThreads), 0, stream, ix, bt, wk, blocks);
        return true;
    case 8192:
        hipLaunchKernelGGL(KScore<8192>, dim3(grid), dim3(kThreads), 0, stream, ix, bt, wk, blocks);
        return true;
    default:
        return false;
    }
}

// Resident scoring workgroups per CU for a block size (registers and LDS), or 0.
int ScoreBlocksPerCu(int blockDocs)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (blockDocs) {
    case 2048:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KScore<2048>, kThreads, 0);
        break;
    case 4096:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KScore<4096>, kThreads, 0);
        break;
    case 8192:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KScore<8192>, kThreads, 0);
        break;
    default:
        break;
    }
    return e == hipSuccess ? n : 0;
}

void LaunchFinal(const BatchView &bt, const WorkView &wk, int32_t maxQ, unsigned long long *hostStats,
                 hipStream_t stream)
{
    hipLaunchKernelGGL(KFinal, dim3(maxQ), dim3(kFinalThreads), 0, stream, bt, wk, hostStats);
}

} // namespace bm25mi
