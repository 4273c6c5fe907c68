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

template <int B>
__global__ __launch_bounds__(kThreads) void KScore(IndexView ix, BatchView bt, WorkView wk, int32_t blocks,
                                                    int32_t granulesPerBlock)
{
    constexpr int kRows = B / (4 * kThreads); // float4 rows per thread
    __shared__ __align__(16) float acc[B];
    __shared__ uint32_t sLo[kMaxTerms], sHi[kMaxTerms];
    __shared__ float sUb[kMaxTerms];
    __shared__ uint32_t sRed[kWaves];
    __shared__ uint32_t sShared[2];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t nq = bt.header[0];
    const int64_t total = static_cast<int64_t>(nq) * blocks;
    unsigned long long scanned = 0, scored = 0, skipped = 0, emitted = 0, overflow = 0;

    for (int64_t task = blockIdx.x; task < total; task += gridDim.x) {
        const int q = static_cast<int>(task % nq);
        const int blk = static_cast<int>(task / nq);
        const int32_t base = blk * B;
        const int32_t end = min(base + B, ix.ndocs);
        const int64_t w0 = bt.qoff[q];
        const int n = static_cast<int>(bt.qoff[q + 1] - w0);
        const int32_t k = bt.k[q];
        const float thr = bt.thr[q];
        __syncthreads(); // previous task is done with LDS

        // 1. Posting slice [lo, hi) of every query term inside this block.
        for (int j = wave; j < n; j += kWaves) {
            const int32_t t = bt.terms[w0 + j];
            uint32_t lo = 0, hi = 0;
            float ub = 0.0F;
            if (t >= 0 && t < ix.vocab) {
                const uint32_t t0 = static_cast<uint32_t>(ix.toff[t]), t1 = static_cast<uint32_t>(ix.toff[t + 1]);
                const int32_t sl = ix.slot[t];
                if (sl >= 0) {
                    const uint32_t *d = ix.dir + static_cast<int64_t>(sl) * (ix.granules + 1);
                    const int64_t g0 = static_cast<int64_t>(blk) * granulesPerBlock;
                    const int64_t g1 = min(g0 + granulesPerBlock, ix.granules);
                    lo = t0 + d[g0];
                    hi = t0 + d[g1];
                    if (thr > 0.0F && hi > lo) {
                        float m = 0.0F;
                        for (int64_t g = g0 + lane; g < g1; g += 64)
                            m = fmaxf(m, ix.bmax[static_cast<int64_t>(sl) * ix.granules + g]);
                        ub = __uint_as_float(WaveMax(__float_as_uint(m)));
                    }
                } else {
                    WaveBounds(ix.post, t0, t1, base, end, lo, hi);
                    ub = ix.tmax[t];
                }
            }
            if (lane == 0) {
                sLo[j] = lo;
                sHi[j] = hi;
                sUb[j] = ub;
            }
        }
        float4 *acc4 = reinterpret_cast<float4 *>(acc);
#pragma unroll
        for (int r = 0; r < kRows; ++r)
            acc4[r * kThreads + tid] = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
        if (tid == 0)
            sShared[0] = max(__atomic_load_n(&wk.tau[q], __ATOMIC_RELAXED), 1u);
        __syncthreads();

        uint32_t tau = sShared[0];
        if (thr > 0.0F) {
            // Upper bound of any doc score in the block: block-max impacts summed in term order.
            float bound = 0.0F;
            for (int j = 0; j < n; ++j)
                if (sHi[j] > sLo[j])
                    bound += sUb[j];
            if (bound < thr) {
                ++skipped;
                continue;
            }
            tau = max(tau, __float_as_uint(thr));
        }
        ++scored;

        // 2. Term-ordered FP32 accumulation. Within one term each doc appears once, so
        //    lanes never collide; the barrier orders the terms.
        for (int j = 0; j < n; ++j) {
            const uint32_t lo = sLo[j], hi = sHi[j];
            if (lo >= hi)
                continue;
            if (tid == 0)
                scanned += hi - lo;
            uint32_t p = lo + tid;
            for (; p + 3 * kThreads < hi; p += 4 * kThreads) {
                const uint2 e0 = ix.post[p];
                const uint2 e1 = ix.post[p + kThreads];
                const uint2 e2 = ix.post[p + 2 * kThreads];
                const uint2 e3 = ix.post[p + 3 * kThreads];
                acc[static_cast<int32_t>(e0.x) - base] += __uint_as_float(e0.y);
                acc[static_cast<int32_t>(e1.x) - base] += __uint_as_float(e1.y);
                acc[static_cast<int32_t>(e2.x) - base] += __uint_as_float(e2.y);
                acc[static_cast<int32_t>(e3.x) - base] += __uint_as_float(e3.y);
            }
            for (; p < hi; p += kThreads) {
                const uint2 e = ix.post[p];
                acc[static_cast<int32_t>(e.x) - base] += __uint_as_float(e.y);
            }
            __syncthreads();
        }

        // 3. Block top-K among docs with score bits >= tau (tau >= 1 excludes zeros).
        uint32_t bits[kRows * 4];
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            const float4 v = acc4[r * kThreads + tid];
            bits[4 * r + 0] = __float_as_uint(v.x);
            bits[4 * r + 1] = __float_as_uint(v.y);
            bits[4 * r + 2] = __float_as_uint(v.z);
            bits[4 * r + 3] = __float_as_uint(v.w);
        }
        auto countGe = [&](uint32_t x) {
            uint32_t c = 0;
#pragma unroll
            for (int i = 0; i < kRows * 4; ++i)
                c += bits[i] >= x ? 1u : 0u;
            return BlockSum<kThreads>(c, sRed);
        };
        const uint32_t candidates = countGe(tau);
        if (candidates == 0)
            continue;
        uint32_t cut = tau;          // emit bits > cut, or bits == cut with local index <= tieIdx
        uint32_t tieIdx = B;         // B: every tie qualifies
        bool exact = false;
        if (candidates > static_cast<uint32_t>(k)) {
            // K-th largest score bits by bisection: count(>= lo) >= k > count(>= hi).
            uint32_t mx = 0;
#pragma unroll
            for (int i = 0; i < kRows * 4; ++i)
                mx = max(mx, bits[i]);
            mx = BlockMax<kThreads>(mx, sRed);
            uint32_t lo = tau, hi = mx + 1;
            while (hi - lo > 1) {
                const uint32_t mid = lo + (hi - lo) / 2;
                if (countGe(mid) >= static_cast<uint32_t>(k))
                    lo = mid;
                else
                    hi = mid;
            }
            cut = lo;
            exact = true;
            const uint32_t greater = lo + 1 <= lo ? 0u : countGe(lo + 1);
            const uint32_t need = static_cast<uint32_t>(k) - greater;
            uint32_t ties = 0;
#pragma unroll
            for (int i = 0; i < kRows * 4; ++i)
                ties += bits[i] == cut ? 1u : 0u;
            ties = BlockSum<kThreads>(ties, sRed);
            if (ties > need) {
                // Smallest local index x with #{ties at index <= x} >= need.
                int32_t a = -1, z = B - 1; // count(<= a) < need <= count(<= z)
                while (z - a > 1) {
                    const int32_t mid = a + (z - a) / 2;
                    uint32_t c = 0;
#pragma unroll
                    for (int r = 0; r < kRows; ++r)
#pragma unroll
                        for (int cc = 0; cc < 4; ++cc) {
                            const int32_t e = r * 4 * kThreads + tid * 4 + cc;
                            c += (bits[4 * r + cc] == cut && e <= mid) ? 1u : 0u;
                        }
                    if (BlockSum<kThreads>(c, sRed) >= need)
                        z = mid;
                    else
                        a = mid;
                }
                tieIdx = static_cast<uint32_t>(z);
            }
        }
        // 4. Append the kept docs to the query's pool.
        uint32_t mine = 0;
#pragma unroll
        for (int r = 0; r < kRows; ++r)
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
                const uint32_t e = r * 4 * kThreads + tid * 4 + cc;
                const uint32_t x = bits[4 * r + cc];
                mine += (x > cut || (x == cut && e <= tieIdx)) ? 1u : 0u;
            }
        uint32_t kept = 0;
        const uint32_t before = BlockExclusive<kThreads>(mine, sRed, &kept);
        if (tid == 0) {
            sShared[1] = atomicAdd(&wk.count[q], kept);
            if (exact)
                atomicMax(&wk.tau[q], cut);
        }
        __syncthreads();
        const uint32_t at = sShared[1];
        const int64_t cap = bt.poolOff[q + 1] - bt.poolOff[q];
        unsigned long long *pool = wk.pool + bt.poolOff[q];
        uint32_t slot = at + before;
#pragma unroll
        for (int r = 0; r < kRows; ++r)
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
                const uint32_t e = r * 4 * kThreads + tid * 4 + cc;
                const uint32_t x = bits[4 * r + cc];
                if (x > cut || (x == cut && e <= tieIdx)) {
                    if (static_cast<int64_t>(slot) < cap)
                        pool[slot] = (static_cast<unsigned long long>(x) << 32) |
                                     static_cast<unsigned long long>(0xFFFFFFFFu - (static_cast<uint32_t>(base) + e));
                    else
                        ++overflow;
                    ++slot;
                }
            }
        if (tid == 0)
            emitted += kept;
    }
    if (tid == 0) {
        atomicAdd(&wk.stats[kStatPostings], scanned);
        atomicAdd(&wk.stats[kStatScored], scored);
        atomicAdd(&wk.stats[kStatSkipped], skipped);
        atomicAdd(&wk.stats[kStatCandidates], emitted);
    }
    if (overflow)
        atomicAdd(&wk.stats[kStatOverflow], overflow);
}

__global__ void KReset(BatchView bt, WorkView wk, int32_t maxQ)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < maxQ) {
        wk.count[i] = 0;
        wk.tau[i] = 1;
    }
    if (i < kStats)
        wk.stats[i] = 0;
}

// Bitonic sort of s[0, n) (n a power of two <= 2 * NT * iterations), descending.
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

__global__ __launch_bounds__(kFinalThreads) void KFinal(BatchView bt, WorkView wk, int32_t ndocs)
{
    __shared__ unsigned long long sKeys[kFinalLds];
    __shared__ unsigned long long sSel[1024];
    __shared__ uint32_t sRed[kFinalThreads / 64];
    __shared__ uint32_t sN;
    const int q = blockIdx.x;
    const int32_t nq = bt.header[0];
    if (q >= nq)
        return;
    const int tid = threadIdx.x;
    const uint32_t k = static_cast<uint32_t>(bt.k[q]);
    const int64_t cap = bt.poolOff[q + 1] - bt.poolOff[q];
    const uint32_t filled = __atomic_load_n(&wk.count[q], __ATOMIC_RELAXED);
    const uint32_t p = static_cast<uint32_t>(min(static_cast<int64_t>(filled), cap));
    const unsigned long long *pool = wk.pool + bt.poolOff[q];
    const bool pruned = bt.thr[q] > 0.0F;
    if (pruned && p < k) {
        if (tid == 0)
            wk.status[q] = 1; // fewer than K survivors: the host reruns this query on the plain path
        return;
    }
    const bool inLds = p <= static_cast<uint32_t>(kFinalLds);
    if (inLds)
        for (uint32_t i = tid; i < p; i += kFinalThreads)
            sKeys[i] = pool[i];
    __syncthreads();
    unsigned long long kappa = 0;
    if (p > k) {
        for (int bit = 63; bit >= 0; --bit) {
            const unsigned long long cand = kappa | (1ULL << bit);
            uint32_t c = 0;
            for (uint32_t i = tid; i < p; i += kFinalThreads)
                c += (inLds ? sKeys[i] : pool[i]) >= cand ? 1u : 0u;
            if (BlockSum<kFinalThreads>(c, sRed) >= k)
                kappa = cand;
        }
    }
    const uint32_t m = p < k ? p : k;
    if (tid == 0)
        sN = 0;
    __syncthreads();
    for (uint32_t i = tid; i < p; i += kFinalThreads) {
        const unsigned long long key = inLds ? sKeys[i] : pool[i];
        if (key >= kappa) {
            const uint32_t at = atomicAdd(&sN, 1u);
            if (at < 1024)
                sSel[at] = key;
        }
    }
    __syncthreads();
    int n2 = 1;
    while (n2 < static_cast<int>(m))
        n2 <<= 1;
    for (int i = static_cast<int>(m) + tid; i < n2; i += kFinalThreads)
        sSel[i] = 0;
    __syncthreads();
    BitonicDesc<kFinalThreads>(sSel, n2);
    const int64_t row = bt.outOff[q];
    for (uint32_t r = tid; r < m; r += kFinalThreads) {
        const unsigned long long key = sSel[r];
        wk.scores[row + r] = __uint_as_float(static_cast<uint32_t>(key >> 32));
        wk.ids[row + r] = static_cast<int32_t>(0xFFFFFFFFu - static_cast<uint32_t>(key & 0xFFFFFFFFu));
    }
    if (tid == 0)
        wk.status[q] = 0;
    if (m < k) {
        // Pad with the lowest-ID docs that scored zero (every positive doc is in the pool).
        __syncthreads();
        for (uint32_t r = tid; r < static_cast<uint32_t>(n2); r += kFinalThreads)
            sKeys[r] = r < m ? static_cast<unsigned long long>(0xFFFFFFFFu - static_cast<uint32_t>(sSel[r] & 0xFFFFFFFFu))
                             : ~0ULL;
        __syncthreads();
        // Ascending by doc ID: sort descending on the complemented IDs.
        for (uint32_t r = tid; r < static_cast<uint32_t>(n2); r += kFinalThreads)
            sKeys[r] = ~sKeys[r];
        __syncthreads();
        BitonicDesc<kFinalThreads>(sKeys, n2);
        if (tid == 0) {
            uint32_t j = 0;
            int32_t d = 0;
            for (uint32_t r = m; r < k; ++r) {
                while (j < m && static_cast<int32_t>(~sKeys[j]) == d) {
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

// Copies the device stats into the host-visible result block.
__global__ void KPublish(WorkView wk, unsigned long long *hostStats)
{
    if (threadIdx.x < kStats)
        hostStats[threadIdx.x] = wk.stats[threadIdx.x];
}

void LaunchReset(const BatchView &bt, const WorkView &wk, int32_t maxQ, hipStream_t stream)
{
    const int threads = 256;
    const int grid = (max(maxQ, kStats) + threads - 1) / threads;
    hipLaunchKernelGGL(KReset, dim3(grid), dim3(threads), 0, stream, bt, wk, maxQ);
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

void LaunchFinal(const BatchView &bt, const WorkView &wk, int32_t maxQ, int32_t ndocs, hipStream_t stream)
{
    hipLaunchKernelGGL(KFinal, dim3(maxQ), dim3(kFinalThreads), 0, stream, bt, wk, ndocs);
}

void LaunchPublish(const WorkView &wk, unsigned long long *hostStats, hipStream_t stream)
{
    hipLaunchKernelGGL(KPublish, dim3(1), dim3(64), 0, stream, wk, hostStats);
}

} // namespace bm25mi
