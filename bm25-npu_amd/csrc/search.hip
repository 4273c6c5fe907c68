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

// Pilot sample size per block: K' = min(K, cap / pilot blocks), where cap is the number
// of 64-bit keys the sorting workgroup's accumulator LDS holds.
__device__ __forceinline__ int32_t PilotK(int32_t k, int32_t pilotBlocks, int32_t cap)
{
    return min(k, cap / max(pilotBlocks, 1));
}

// The scoring kernel: persistent workgroups walk (query, block) tasks. Pilot blocks
// (every kPilotStride-th block) come first; each keeps its exact block top-K (and its
// top-K' in the pilot sample); the last pilot task of a query sorts the sample and
// raises the query's lower bound tau to the sample's K-th score. Every later task keeps
// its block top-K among docs scoring at least tau, so few candidates reach the pool.
template <int B>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void KScore(
    IndexView ix, BatchView bt, WorkView wk, int32_t blocks, int32_t granulesPerBlock)
{
    constexpr int kRows = B / (4 * kThreads); // float4 rows per thread
    constexpr int U = kLoadsPerThread;        // 256-posting columns per round
    constexpr int kSampleCap = B / 2; // pilot keys sortable in the accumulator LDS
    static_assert(kSampleCap <= kPilotCap, "pilot pool stride");
    __shared__ __align__(16) float acc[B];
    __shared__ uint32_t sLo[kMaxTerms], sPre[kMaxTerms + 1], sCol[kMaxTerms + 1];
    __shared__ float sUb[kMaxTerms];
    __shared__ uint2 sRound[U];
    __shared__ uint32_t sRed[kWaves];
    __shared__ uint32_t sShared[4];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t nq = bt.header[0];
    const int32_t pilotBlocks = (blocks + kPilotStride - 1) / kPilotStride;
    const int64_t pilotTasks = static_cast<int64_t>(nq) * pilotBlocks;
    const int64_t total = static_cast<int64_t>(nq) * blocks;
    unsigned long long scanned = 0, scored = 0, skipped = 0, emitted = 0, overflow = 0, emitting = 0, exactTasks = 0;

    for (int64_t task = blockIdx.x; task < total; task += gridDim.x) {
        int q, blk;
        const bool isPilot = task < pilotTasks;
        if (isPilot) {
            q = static_cast<int>(task % nq);
            blk = static_cast<int>(task / nq) * kPilotStride;
        } else {
            const int64_t t = task - pilotTasks;
            q = static_cast<int>(t % nq);
            const int64_t j = t / nq; // j-th block that is not a pilot block
            blk = static_cast<int>(kPilotStride * (j / (kPilotStride - 1)) + j % (kPilotStride - 1) + 1);
        }
        const float thr = bt.thr[q];
        const bool samples = isPilot && thr <= 0.0F; // a predicted threshold replaces the sample
        const int32_t base = blk * B;
        const int32_t end = min(base + B, ix.ndocs);
        const int64_t w0 = bt.qoff[q];
        const int n = static_cast<int>(bt.qoff[q + 1] - w0);
        const int32_t k = bt.k[q];
        __syncthreads(); // previous task is done with LDS

        // 1. Posting slice of every query word inside this block: block directory for
        //    long lists, a two-round 64-ary wave search for short ones.
        for (int j = wave; j < n; j += kWaves) {
            const uint4 info = bt.words[w0 + j];
            uint32_t lo = 0, hi = 0;
            float ub = 0.0F;
            if (info.y > 0) {
                const int32_t sl = static_cast<int32_t>(info.z);
                if (sl >= 0) {
                    const uint32_t *d = ix.dir + static_cast<int64_t>(sl) * (ix.granules + 1);
                    const int64_t g0 = static_cast<int64_t>(blk) * granulesPerBlock;
                    const int64_t g1 = min(g0 + granulesPerBlock, ix.granules);
                    lo = info.x + d[g0];
                    hi = info.x + d[g1];
                    if (thr > 0.0F && hi > lo) {
                        float m = 0.0F;
                        for (int64_t g = g0 + lane; g < g1; g += 64)
                            m = fmaxf(m, ix.bmax[static_cast<int64_t>(sl) * ix.granules + g]);
                        ub = __uint_as_float(WaveMax(__float_as_uint(m)));
                    }
                } else {
                    WaveBounds(ix.post, info.x, info.x + info.y, base, end, lo, hi);
                    ub = __uint_as_float(info.w);
                }
            }
            if (lane == 0) {
                sLo[j] = lo;
                sPre[j + 1] = hi - lo; // length; turned into a prefix below
                sUb[j] = ub;
            }
        }
        float4 *acc4 = reinterpret_cast<float4 *>(acc);
#pragma unroll
        for (int r = 0; r < kRows; ++r)
            acc4[r * kThreads + tid] = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
        __syncthreads();
        if (wave == 0) {
            // Prefixes of slice lengths and of 256-posting column counts (one lane per word).
            const uint32_t len = lane < n ? sPre[lane + 1] : 0u;
            uint32_t x = len, c = (len + kThreads - 1) / kThreads;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o), z = __shfl_up(c, o);
                if (lane >= o) {
                    x += y;
                    c += z;
                }
            }
            if (lane < n) {
                sPre[lane + 1] = x;
                sCol[lane + 1] = c;
            }
            if (lane == 0) {
                sPre[0] = 0;
                sCol[0] = 0;
                sShared[0] = max(__atomic_load_n(&wk.tau[q], __ATOMIC_RELAXED), 1u);
                sShared[2] = 1;
            }
        }
        __syncthreads();
        if (thr > 0.0F && tid == 0) {
            // Upper bound of any score in the block: per-word maxima summed in word order.
            float bound = 0.0F;
            for (int j = 0; j < n; ++j)
                if (sPre[j + 1] > sPre[j])
                    bound += sUb[j];
            sShared[2] = bound < thr ? 0u : 1u;
        }
        if (thr > 0.0F)
            __syncthreads();
        const bool skip = sShared[2] == 0;
        uint32_t tau = samples ? 1u : sShared[0];
        if (thr > 0.0F)
            tau = max(tau, __float_as_uint(thr));
        if (skip) {
            ++skipped;
        } else {
            ++scored;
            // 2. Term-ordered FP32 accumulation. Each word's slice is cut into 256-posting
            //    columns; a round issues up to U columns' loads (one 8-byte load per thread
            //    per column) before any add, then adds them column by column with a barrier
            //    where the word changes. Within one word each doc appears once, so lanes never
            //    collide, and every doc's score is summed in query-word order.
            const uint32_t columns = sCol[n];
            if (tid == 0)
                scanned += sPre[n];
            for (uint32_t r0 = 0; r0 < columns; r0 += U) {
                if (wave == 0 && lane < U) {
                    const uint32_t col = r0 + lane;
                    uint2 entry = make_uint2(0u, 0xFFFF0000u);
                    if (col < columns) {
                        int j = 0;
                        while (sCol[j + 1] <= col)
                            ++j;
                        const uint32_t i = col - sCol[j];
                        const uint32_t len = sPre[j + 1] - sPre[j];
                        entry = make_uint2(sLo[j] + i * kThreads,
                                           min(static_cast<uint32_t>(kThreads), len - i * kThreads) |
                                               (static_cast<uint32_t>(j) << 16));
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
                // Unused columns carry word 0xFFFF and count 0: no barrier, no add. The
                // barrier condition reads LDS values shared by the whole workgroup.
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

            // 3. Block selection. Scores are re-read from LDS in each pass (8 x 16-byte
            //    reads per thread) so they do not occupy registers across the passes.
            const uint4 *bits4 = reinterpret_cast<const uint4 *>(acc);
            auto countIf = [&](auto pred) {
                uint32_t c = 0;
#pragma unroll
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
            // Cut for the block's top-kk among docs with bits >= floor: keep bits > cut, or
            // bits == cut at local index <= tieIdx. Returns false when nothing qualifies.
            auto selectCut = [&](uint32_t kk, uint32_t floor, uint32_t &cut, uint32_t &tieIdx, bool &exact) {
                const uint32_t candidates = countGe(floor);
                cut = floor;
                tieIdx = B;
                exact = false;
                if (candidates == 0)
                    return false;
                if (candidates <= kk)
                    return true;
                // kk-th largest score bits by bisection: count(>= lo) >= kk > count(>= hi).
                uint32_t mx = 0;
#pragma unroll
                for (int r = 0; r < kRows; ++r) {
                    const uint4 v = bits4[r * kThreads + tid];
                    mx = max(mx, max(max(v.x, v.y), max(v.z, v.w)));
                }
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
                exact = true;
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
                return true;
            };
            // Appends the kept docs to pool[0, cap) through *counter; returns how many.
            auto emit = [&](uint32_t cut, uint32_t tieIdx, unsigned long long *pool, uint32_t *counter, int64_t cap) {
                auto keepIt = [cut, tieIdx](uint32_t b, uint32_t e) { return b > cut || (b == cut && e <= tieIdx); };
                const uint32_t mine = countIf(keepIt);
                uint32_t kept = 0;
                const uint32_t before = BlockExclusive<kThreads>(mine, sRed, &kept);
                if (tid == 0)
                    sShared[1] = atomicAdd(counter, kept);
                __syncthreads();
                if (mine > 0) {
                    uint32_t slot = sShared[1] + before;
#pragma unroll
                    for (int r = 0; r < kRows; ++r) {
                        const uint4 v = bits4[r * kThreads + tid];
                        const uint32_t vals[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                        for (int cc = 0; cc < 4; ++cc) {
                            const uint32_t e = r * 4 * kThreads + tid * 4 + cc;
                            if (keepIt(vals[cc], e)) {
                                if (static_cast<int64_t>(slot) < cap)
                                    pool[slot] = (static_cast<unsigned long long>(vals[cc]) << 32) |
                                                 static_cast<unsigned long long>(0xFFFFFFFFu -
                                                                                 (static_cast<uint32_t>(base) + e));
                                else
                                    ++overflow;
                                ++slot;
                            }
                        }
                    }
                }
                return kept;
            };

            uint32_t cut, tieIdx;
            bool exact;
            if (selectCut(static_cast<uint32_t>(k), tau, cut, tieIdx, exact)) {
                const uint32_t kept =
                    emit(cut, tieIdx, wk.pool + bt.poolOff[q], &wk.count[q], bt.poolOff[q + 1] - bt.poolOff[q]);
                if (exact) {
                    ++exactTasks;
                    if (tid == 0)
                        atomicMax(&wk.tau[q], cut); // the block's K-th score bounds the query's
                }
                if (tid == 0) {
                    emitted += kept;
                    ++emitting;
                }
            }
            if (samples) {
                // The sample keeps the block's top-K' (the same cut as above when K' == K).
                const int32_t kp = PilotK(k, pilotBlocks, kSampleCap);
                if (kp > 0) {
                    if (kp < k)
                        selectCut(static_cast<uint32_t>(kp), 1u, cut, tieIdx, exact);
                    emit(cut, tieIdx, wk.pilot + static_cast<int64_t>(q) * kPilotCap, &wk.pilotCount[q],
                         static_cast<int64_t>(pilotBlocks) * kp);
                }
            }
        }
        if (samples) {
            // Publish this pilot task; the last one of the query turns the sample into tau.
            __threadfence();
            __syncthreads();
            if (tid == 0)
                sShared[3] = atomicAdd(&wk.pilotDone[q], 1u);
            __syncthreads();
            if (sShared[3] == static_cast<uint32_t>(pilotBlocks - 1)) {
                __threadfence();
                const int32_t kp = PilotK(k, pilotBlocks, kSampleCap);
                const uint32_t p = min(__atomic_load_n(&wk.pilotCount[q], __ATOMIC_RELAXED),
                                       static_cast<uint32_t>(pilotBlocks * kp));
                if (p >= static_cast<uint32_t>(k)) {
                    unsigned long long *sorted = reinterpret_cast<unsigned long long *>(acc);
                    const unsigned long long *sample = wk.pilot + static_cast<int64_t>(q) * kPilotCap;
                    int n2 = 1;
                    while (n2 < static_cast<int>(p))
                        n2 <<= 1;
                    __syncthreads();
                    for (int i = tid; i < n2; i += kThreads)
                        sorted[i] = i < static_cast<int>(p) ? sample[i] : 0ULL;
                    __syncthreads();
                    BitonicDesc<kThreads>(sorted, n2);
                    if (tid == 0)
                        atomicMax(&wk.tau[q], max(1u, static_cast<uint32_t>(sorted[k - 1] >> 32)));
                }
                if (tid == 0) {
                    wk.pilotCount[q] = 0;
                    wk.pilotDone[q] = 0;
                }
            }
        }
    }
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
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < maxQ) {
        wk.count[i] = 0;
        wk.pilotCount[i] = 0;
        wk.pilotDone[i] = 0;
        wk.tau[i] = 1;
    }
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
    const uint32_t tau = wk.tau[q];
    const unsigned long long *pool = wk.pool + bt.poolOff[q];
    const bool pruned = bt.thr[q] > 0.0F;
    if (tid == 0)
        sN = 0;
    __syncthreads();
    if (tid == 0) {
        // Restore the per-query counters for the next Search.
        wk.count[q] = 0;
        wk.tau[q] = 1;
    }
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
    const int grid = (max(maxQ, kStats) + threads - 1) / threads;
    hipLaunchKernelGGL(KReset, dim3(grid), dim3(threads), 0, stream, wk, maxQ);
}

bool LaunchScore(int blockDocs, const IndexView &ix, const BatchView &bt, const WorkView &wk, int32_t blocks,
                 int32_t granulesPerBlock, int32_t grid, hipStream_t stream)
{
    switch (blockDocs) {
    case 2048:
        hipLaunchKernelGGL(KScore<2048>, dim3(grid), dim3(kThreads), 0, stream, ix, bt, wk, blocks, granulesPerBlock);
        return true;
    case 4096:
        hipLaunchKernelGGL(KScore<4096>, dim3(grid), dim3(kThreads), 0, stream, ix, bt, wk, blocks, granulesPerBlock);
        return true;
    case 8192:
        hipLaunchKernelGGL(KScore<8192>, dim3(grid), dim3(kThreads), 0, stream, ix, bt, wk, blocks, granulesPerBlock);
        return true;
    default:
        return false;
    }
}

void LaunchFinal(const BatchView &bt, const WorkView &wk, int32_t maxQ, unsigned long long *hostStats,
                 hipStream_t stream)
{
    hipLaunchKernelGGL(KFinal, dim3(maxQ), dim3(kFinalThreads), 0, stream, bt, wk, hostStats);
}

} // namespace bm25mi
