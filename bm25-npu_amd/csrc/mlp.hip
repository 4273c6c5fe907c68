// Training and inference kernels of the learned K-th-score model (no reference
// counterpart): an MLP in -> 64 -> 64 -> 1 with ReLU, FP32 throughout.
//
// Parameter layout (flat): W1[64][in] b1[64] W2[64][64] b2[64] W3[64] b3[1].
//   forward   one thread per sample, weights staged in LDS (broadcast reads).
//   backward  each workgroup walks 64-sample tiles: the tile's forward pass into LDS,
//             the mean-squared-error gradient, and its products on the matrix cores
//             (MFMA 16x16x4 FP32), with the weight gradients accumulated in the
//             waves' MFMA accumulators across tiles; partials go to a
//             per-workgroup buffer and a second kernel sums them in a fixed
//             order, so gradients are deterministic run to run.
//   adam      elementwise Adam update with bias correction; a training step sums the
//             partials and applies the update in one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"

namespace bm25mi {

constexpr int kH = 64;
constexpr int kMaxIn = 32;
constexpr int kTile = 64;
constexpr int kMlpThreads = 256;
constexpr int kBackwardGrid = 512;

__host__ __device__ inline int64_t MlpParams(int in) { return int64_t(kH) * in + kH + kH * kH + kH + kH + 1; }

__global__ __launch_bounds__(kMlpThreads) void KMlpForward(const float *__restrict__ p, const float *__restrict__ x,
                                                           int n, int in, float *__restrict__ y)
{
    __shared__ float sW1[kH * kMaxIn], sB1[kH], sW2[kH * kH], sB2[kH], sW3[kH];
    const float *W1 = p, *B1 = W1 + kH * in, *W2 = B1 + kH, *B2 = W2 + kH * kH, *W3 = B2 + kH, *B3 = W3 + kH;
    for (int i = threadIdx.x; i < kH * in; i += blockDim.x)
        sW1[i] = W1[i];
    for (int i = threadIdx.x; i < kH * kH; i += blockDim.x)
        sW2[i] = W2[i];
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
        sB1[i] = B1[i];
        sB2[i] = B2[i];
        sW3[i] = W3[i];
    }
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    float xv[kMaxIn];
    for (int i = 0; i < in; ++i)
        xv[i] = x[int64_t(s) * in + i];
    float h1[kH];
#pragma unroll
    for (int j = 0; j < kH; ++j) {
        float z = sB1[j];
        for (int i = 0; i < in; ++i)
            z = fmaf(sW1[j * in + i], xv[i], z);
        h1[j] = fmaxf(z, 0.0F);
    }
    float out = B3[0];
    for (int k = 0; k < kH; ++k) {
        float z = sB2[k];
#pragma unroll
        for (int j = 0; j < kH; ++j)
            z = fmaf(sW2[k * kH + j], h1[j], z);
        out = fmaf(sW3[k], fmaxf(z, 0.0F), out);
    }
    y[s] = out;
}

typedef float MfmaF4 __attribute__((ext_vector_type(4)));

// LDS row strides of the backward kernel's tiles: odd, so that the 16 rows one operand fetch
// spans fall in distinct banks.
constexpr int kXS = kMaxIn + 1;
constexpr int kHS = kH + 1;

// acc[t] += A[m0, m0+16) x [0, kK) * B[0, kK) x [16t, 16t+16) on the matrix cores
// (v_mfma_f32_16x16x4f32: lane l gives A[l % 16][l / 16] and B[l / 16][l % 16] per step of 4,
// and holds C[4 (l / 16) + r][l % 16] in element r). a points at A[m0][0], b at B[0][0];
// A[m][k] = a[m * kAS + k * kAK], B[k][n] = b[k * kBS + n * kBN].
template <int kTiles, int kK, int kAS, int kAK, int kBS, int kBN>
__device__ inline void MfmaRows(MfmaF4 (&acc)[kTiles], const float *a, const float *b, int lane)
{
    const int r = lane & 15, c = lane >> 4;
    a += r * kAS + c * kAK;
    b += c * kBS + r * kBN;
#pragma unroll
    for (int k0 = 0; k0 < kK; k0 += 4) {
        const float av = a[k0 * kAK];
#pragma unroll
        for (int t = 0; t < kTiles; ++t)
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b[k0 * kBS + 16 * t * kBN], acc[t], 0, 0, 0);
    }
}

// Partial gradients of one workgroup: partial[blockIdx.x][0, P) and lossPartial[blockIdx.x].
// Each 64-sample tile runs its five 64-wide products on the matrix cores, wave w taking the
// output rows 16w .. 16w+15: h1 = relu(x W1^T + b1), h2 = relu(h1 W2^T + b2), then
// dW2 += dz2^T h1, dz1 = (dz2 W2) * (h1 > 0) and dW1 += dz1^T x, with dz2 = dy w3 * (h2 > 0).
// The weight gradients stay in the waves' accumulators across tiles; the 64-long sums (y, the
// biases' and W3's gradients) are per-lane loops of wave 0.
__global__ __launch_bounds__(kMlpThreads) void KMlpBackward(const float *__restrict__ p, const float *__restrict__ x,
                                                            const float *__restrict__ target, int n, int in,
                                                            float *__restrict__ partial,
                                                            float *__restrict__ lossPartial)
{
    __shared__ float sX[kTile * kXS];  // x tile; columns in .. kMaxIn-1 and rows past the batch are 0
    __shared__ float sW1[kH * kXS];    // W1; columns in .. kMaxIn-1 are 0
    __shared__ float sW2[kH * kHS];
    __shared__ float sH1[kTile * kHS]; // h1, later dz1
    __shared__ float sH2[kTile * kHS]; // h2, later dz2
    __shared__ float sDy[kTile];
    __shared__ float sRed[kMlpThreads / 64];
    const float *W1 = p, *B1 = W1 + kH * in, *W2 = B1 + kH, *B2 = W2 + kH * kH, *W3 = B2 + kH, *B3 = W3 + kH;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < kH * kMaxIn; i += kMlpThreads) {
        const int j = i / kMaxIn, c = i - j * kMaxIn;
        sW1[j * kXS + c] = c < in ? W1[j * in + c] : 0.0F;
    }
    for (int i = tid; i < kH * kH; i += kMlpThreads)
        sW2[(i / kH) * kHS + i % kH] = W2[i];
    // Lane's place in a 16 x 16 output tile: rows 4 * quad + r, column col.
    const int col = lane & 15, quad = lane >> 4, m0 = 16 * wave;
    float b1v[4], b2v[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        b1v[t] = B1[16 * t + col];
        b2v[t] = B2[16 * t + col];
    }
    MfmaF4 gW2[4], gW1[kMaxIn / 16]; // dW2[m0 + 4 quad + r][16t + col], dW1[m0 + 4 quad + r][16t + col]
#pragma unroll
    for (int t = 0; t < 4; ++t)
        gW2[t] = MfmaF4{0.0F, 0.0F, 0.0F, 0.0F};
#pragma unroll
    for (int t = 0; t < kMaxIn / 16; ++t)
        gW1[t] = MfmaF4{0.0F, 0.0F, 0.0F, 0.0F};
    float gB1 = 0, gB2 = 0, gW3 = 0, gB3 = 0, loss = 0; // wave 0: unit = lane
    const float w3 = W3[lane];
    const float scale = 2.0F / static_cast<float>(n);
    __syncthreads();
    for (int64_t t0 = int64_t(blockIdx.x) * kTile; t0 < n; t0 += int64_t(gridDim.x) * kTile) {
        const int rows = static_cast<int>(min<int64_t>(kTile, n - t0));
        for (int i = tid; i < kTile * kMaxIn; i += kMlpThreads) {
            const int s = i / kMaxIn, c = i - s * kMaxIn;
            sX[s * kXS + c] = s < rows && c < in ? x[(t0 + s) * in + c] : 0.0F;
        }
        __syncthreads();
        {
            MfmaF4 acc[4] = {};
            MfmaRows<4, kMaxIn, kXS, 1, 1, kXS>(acc, sX + m0 * kXS, sW1, lane);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    sH1[(m0 + 4 * quad + r) * kHS + 16 * t + col] = fmaxf(acc[t][r] + b1v[t], 0.0F);
        }
        __syncthreads();
        {
            MfmaF4 acc[4] = {};
            MfmaRows<4, kH, kHS, 1, 1, kHS>(acc, sH1 + m0 * kHS, sW2, lane);
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    sH2[(m0 + 4 * quad + r) * kHS + 16 * t + col] = fmaxf(acc[t][r] + b2v[t], 0.0F);
        }
        __syncthreads();
        // y, dy = 2 (y - target) / n, then dW3, db3 (wave 0, sample / unit = lane).
        if (wave == 0) {
            float yv = B3[0];
            for (int k = 0; k < kH; ++k)
                yv = fmaf(W3[k], sH2[lane * kHS + k], yv);
            const float d = lane < rows ? yv - target[t0 + lane] : 0.0F;
            loss = fmaf(d, d, loss);
            sDy[lane] = d * scale;
        }
        __syncthreads();
        if (wave == 0) {
            for (int s = 0; s < kTile; ++s)
                gW3 = fmaf(sDy[s], sH2[s * kHS + lane], gW3);
            if (lane == 0)
                for (int s = 0; s < kTile; ++s)
                    gB3 += sDy[s];
        }
        __syncthreads();
        // dz2 = dy * W3 * (h2 > 0) in place of h2.
        for (int s = wave; s < kTile; s += 4)
            sH2[s * kHS + lane] = sH2[s * kHS + lane] > 0.0F ? sDy[s] * w3 : 0.0F;
        __syncthreads();
        // dW2[k][j] += sum_s dz2[s][k] h1[s][j] (rows k), and dz2 W2 (rows s) before h1 is overwritten.
        MfmaRows<4, kTile, 1, kHS, kHS, 1>(gW2, sH2 + m0, sH1, lane);
        MfmaF4 dz[4] = {};
        MfmaRows<4, kH, kHS, 1, kHS, 1>(dz, sH2 + m0 * kHS, sW2, lane);
        if (wave == 0)
            for (int s = 0; s < kTile; ++s)
                gB2 += sH2[s * kHS + lane];
        __syncthreads();
        // dz1 = (dz2 W2) * (h1 > 0) in place of h1: each element read and written by its own lane.
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float &h = sH1[(m0 + 4 * quad + r) * kHS + 16 * t + col];
                h = h > 0.0F ? dz[t][r] : 0.0F;
            }
        __syncthreads();
        // dW1[j][i] += sum_s dz1[s][j] x[s][i] (rows j, columns i < kMaxIn).
        MfmaRows<kMaxIn / 16, kTile, 1, kHS, kXS, 1>(gW1, sH1 + m0, sX, lane);
        if (wave == 0)
            for (int s = 0; s < kTile; ++s)
                gB1 += sH1[s * kHS + lane];
        __syncthreads();
    }
    float *out = partial + int64_t(blockIdx.x) * MlpParams(in);
    float *oW1 = out, *oB1 = oW1 + kH * in, *oW2 = oB1 + kH, *oB2 = oW2 + kH * kH, *oW3 = oB2 + kH, *oB3 = oW3 + kH;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            oW2[(m0 + 4 * quad + r) * kH + 16 * t + col] = gW2[t][r];
#pragma unroll
    for (int t = 0; t < kMaxIn / 16; ++t)
        if (16 * t + col < in)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                oW1[(m0 + 4 * quad + r) * in + 16 * t + col] = gW1[t][r];
    if (wave == 0) {
        oB1[lane] = gB1;
        oB2[lane] = gB2;
        oW3[lane] = gW3;
        if (lane == 0)
            oB3[0] = gB3;
    }
    // Loss partial: fixed-order block sum.
    for (int o = 32; o > 0; o >>= 1)
        loss += __shfl_xor(loss, o);
    if (lane == 0)
        sRed[wave] = loss;
    __syncthreads();
    if (tid == 0)
        lossPartial[blockIdx.x] = sRed[0] + sRed[1] + sRed[2] + sRed[3];
}

// Fixed-order sum of the workgroup partials of parameter i.
__device__ inline float PartialSum(const float *__restrict__ partial, int blocks, int64_t count, int64_t i)
{
    float s = 0.0F;
    for (int b = 0; b < blocks; ++b)
        s += partial[int64_t(b) * count + i];
    return s;
}

__device__ inline float MeanLoss(const float *__restrict__ lossPartial, int blocks, int n)
{
    float s = 0.0F;
    for (int b = 0; b < blocks; ++b)
        s += lossPartial[b];
    return s / static_cast<float>(n);
}

// Adam update of parameter i with gradient gi; c1, c2 are the bias corrections 1 - beta^step.
__device__ inline void AdamAt(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v, int64_t i, float gi,
                              float lr, float beta1, float beta2, float eps, float c1, float c2)
{
    const float mi = beta1 * m[i] + (1.0F - beta1) * gi;
    const float vi = beta2 * v[i] + (1.0F - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// grad += the summed partials, lossOut[0] = mean loss.
__global__ void KMlpReduce(const float *__restrict__ partial, const float *__restrict__ lossPartial, int blocks,
                           int64_t count, int n, float *__restrict__ grad, float *__restrict__ lossOut)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count)
        grad[i] += PartialSum(partial, blocks, count, i);
    if (i == 0 && lossOut)
        lossOut[0] = MeanLoss(lossPartial, blocks, n);
}

__global__ void KMlpAdam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                         float *__restrict__ v, int64_t count, float lr, float beta1, float beta2, float eps,
                         float c1, float c2)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count)
        AdamAt(p, m, v, i, g[i], lr, beta1, beta2, eps, c1, c2);
}

// One training step after the backward kernel, in one launch: each parameter's gradient is the
// fixed-order sum of the workgroup partials, followed by its Adam update; lossOut[0] += the
// minibatch's mean loss (the epoch's running sum).
__global__ void KMlpReduceAdam(const float *__restrict__ partial, const float *__restrict__ lossPartial, int blocks,
                               int64_t count, int n, float *__restrict__ p, float *__restrict__ m,
                               float *__restrict__ v, float *__restrict__ lossOut, float lr, float beta1, float beta2,
                               float eps, float c1, float c2)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count)
        AdamAt(p, m, v, i, PartialSum(partial, blocks, count, i), lr, beta1, beta2, eps, c1, c2);
    if (i == 0)
        lossOut[0] += MeanLoss(lossPartial, blocks, n);
}

// Workgroup partials of bm25mi_mlp_backward, shared by its calls: each call records scratchDone
// after its last kernel, and the next call's stream waits on it (or the host does, before the
// buffer grows), so calls on any streams take the buffer in turn without a host wait per call.
namespace {
std::mutex scratchMutex;
float *scratch = nullptr;
size_t scratchFloats = 0;
hipEvent_t scratchDone = nullptr;
} // namespace

} // namespace bm25mi

using namespace bm25mi;

#define BM25MI_MLP_HIP(expr, what)                                                                                     \
    do {                                                                                                               \
        const hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess)                                                                                          \
            throw DeviceError(static_cast<int64_t>(e_), std::string(what) + ": " + hipGetErrorString(e_));         \
    } while (0)

extern "C" int64_t bm25mi_mlp_param_count(int32_t in_features) { return MlpParams(in_features); }

extern "C" int bm25mi_mlp_forward(const float *params, const float *x, int32_t n, int32_t in_features, float *y,
                                  void *stream)
{
    return Invoke([&] {
        Require(params && x && y && n >= 1 && in_features >= 1 && in_features <= kMaxIn, "Invalid MLP forward");
        const int grid = (n + kMlpThreads - 1) / kMlpThreads;
        hipLaunchKernelGGL(KMlpForward, dim3(grid), dim3(kMlpThreads), 0, static_cast<hipStream_t>(stream), params,
                           x, n, in_features, y);
        BM25MI_MLP_HIP(hipGetLastError(), "MLP forward launch");
    });
}

extern "C" int bm25mi_mlp_backward(const float *params, const float *x, const float *target, int32_t n,
                                   int32_t in_features, float *grad, float *loss_out, void *stream)
{
    return Invoke([&] {
        Require(params && x && target && grad && n >= 1 && in_features >= 1 && in_features <= kMaxIn,
                "Invalid MLP backward");
        const int64_t count = MlpParams(in_features);
        const int blocks = static_cast<int>(std::min<int64_t>(kBackwardGrid, (int64_t(n) + kTile - 1) / kTile));
        std::lock_guard<std::mutex> guard(scratchMutex);
        auto s = static_cast<hipStream_t>(stream);
        if (!scratchDone)
            BM25MI_MLP_HIP(hipEventCreateWithFlags(&scratchDone, hipEventDisableTiming), "MLP scratch event");
        const size_t need = static_cast<size_t>(blocks) * (count + 1);
        if (need > scratchFloats) {
            BM25MI_MLP_HIP(hipEventSynchronize(scratchDone), "MLP scratch sync");
            if (scratch)
                BM25MI_MLP_HIP(hipFree(scratch), "MLP scratch free");
            scratch = nullptr;
            scratchFloats = 0;
            BM25MI_MLP_HIP(hipMalloc(&scratch, need * sizeof(float)), "MLP scratch alloc");
            scratchFloats = need;
        } else {
            BM25MI_MLP_HIP(hipStreamWaitEvent(s, scratchDone, 0), "MLP scratch wait");
        }
        float *partial = scratch, *lossPartial = scratch + static_cast<size_t>(blocks) * count;
        hipLaunchKernelGGL(KMlpBackward, dim3(blocks), dim3(kMlpThreads), 0, s, params, x, target, n, in_features,
                           partial, lossPartial);
        BM25MI_MLP_HIP(hipGetLastError(), "MLP backward launch");
        const int rgrid = static_cast<int>((count + 255) / 256);
        hipLaunchKernelGGL(KMlpReduce, dim3(rgrid), dim3(256), 0, s, partial, lossPartial, blocks, count, n, grad,
                           loss_out);
        BM25MI_MLP_HIP(hipGetLastError(), "MLP reduce launch");
        BM25MI_MLP_HIP(hipEventRecord(scratchDone, s), "MLP scratch record");
    });
}

extern "C" int bm25mi_mlp_adam(float *params, const float *grad, float *m, float *v, int64_t count, float lr,
                               float beta1, float beta2, float eps, int32_t step, void *stream)
{
    return Invoke([&] {
        Require(params && grad && m && v && count >= 1 && step >= 1, "Invalid Adam step");
        const float c1 = 1.0F - std::pow(beta1, static_cast<float>(step));
        const float c2 = 1.0F - std::pow(beta2, static_cast<float>(step));
        const int grid = static_cast<int>((count + 255) / 256);
        hipLaunchKernelGGL(KMlpAdam, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream), params, grad, m, v,
                           count, lr, beta1, beta2, eps, c1, c2);
        BM25MI_MLP_HIP(hipGetLastError(), "Adam launch");
    });
}

// ---------------------------------------------------------------------------------------
// Host-pointer entry points: device buffers are owned here, so callers need no GPU
// allocator of their own.
namespace {

struct DeviceBuffer {
    float *p = nullptr;
    explicit DeviceBuffer(size_t n)
    {
        BM25MI_MLP_HIP(hipMalloc(&p, std::max<size_t>(1, n) * sizeof(float)), "MLP buffer alloc");
    }
    ~DeviceBuffer()
    {
        if (p)
            (void)hipFree(p);
    }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    void Upload(const float *src, size_t n, hipStream_t s)
    {
        BM25MI_MLP_HIP(hipMemcpyAsync(p, src, n * sizeof(float), hipMemcpyHostToDevice, s), "MLP upload");
    }
    void Download(float *dst, size_t n, hipStream_t s)
    {
        BM25MI_MLP_HIP(hipMemcpyAsync(dst, p, n * sizeof(float), hipMemcpyDeviceToHost, s), "MLP download");
    }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { BM25MI_MLP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "MLP stream"); }
    ~Stream()
    {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

inline uint64_t SplitMix64(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline double Uniform01(uint64_t &x) { return static_cast<double>(SplitMix64(x) >> 11) * (1.0 / 9007199254740992.0); }

void CheckShapes(const float *x, const float *y, int64_t n, int32_t in)
{
    Require(x && y && n >= 1 && n <= (int64_t(1) << 30) && in >= 1 && in <= kMaxIn, "Invalid MLP data shape");
}

} // namespace

// He-uniform initial weights (bound sqrt(6 / fan_in)), zero biases, from a seeded stream.
extern "C" int bm25mi_mlp_init(int32_t in_features, uint64_t seed, float *params_out)
{
    return Invoke([&] {
        Require(params_out && in_features >= 1 && in_features <= kMaxIn, "Invalid MLP init");
        const int in = in_features;
        float *W1 = params_out, *B1 = W1 + kH * in, *W2 = B1 + kH, *B2 = W2 + kH * kH, *W3 = B2 + kH, *B3 = W3 + kH;
        uint64_t st = seed;
        auto fill = [&](float *w, int count, int fanIn) {
            const double bound = std::sqrt(6.0 / fanIn);
            for (int i = 0; i < count; ++i)
                w[i] = static_cast<float>((2.0 * Uniform01(st) - 1.0) * bound);
        };
        fill(W1, kH * in, in);
        fill(W2, kH * kH, kH);
        fill(W3, kH, kH);
        std::fill(B1, B1 + kH, 0.0F);
        std::fill(B2, B2 + kH, 0.0F);
        B3[0] = 0.0F;
    });
}

// Mean-squared-error gradient of params on (x, target), computed on the GPU, returned to host.
extern "C" int bm25mi_mlp_gradient(const float *params, const float *x, const float *target, int32_t n,
                                   int32_t in_features, float *grad_out, float *loss_out)
{
    return Invoke([&] {
        CheckShapes(x, target, n, in_features);
        Require(params && grad_out, "Null MLP gradient arguments");
        const int64_t count = MlpParams(in_features);
        Stream st;
        DeviceBuffer dp(count), dx(static_cast<size_t>(n) * in_features), dt(n), dg(count), dl(1);
        dp.Upload(params, count, st.s);
        dx.Upload(x, static_cast<size_t>(n) * in_features, st.s);
        dt.Upload(target, n, st.s);
        BM25MI_MLP_HIP(hipMemsetAsync(dg.p, 0, count * sizeof(float), st.s), "MLP gradient clear");
        const int rc = bm25mi_mlp_backward(dp.p, dx.p, dt.p, n, in_features, dg.p, dl.p, st.s);
        if (rc != BM25MI_OK)
            throw DeviceError(0, "MLP backward failed");
        dg.Download(grad_out, count, st.s);
        float loss = 0;
        dl.Download(&loss, 1, st.s);
        BM25MI_MLP_HIP(hipStreamSynchronize(st.s), "MLP gradient sync");
        if (loss_out)
            *loss_out = loss;
    });
}

// y[n] = MLP(x[n, in]) on the GPU, host arrays in and out.
extern "C" int bm25mi_mlp_predict(const float *params, const float *x, int32_t n, int32_t in_features, float *y_out)
{
    return Invoke([&] {
        CheckShapes(x, y_out, n, in_features);
        Require(params, "Null MLP parameters");
        const int64_t count = MlpParams(in_features);
        Stream st;
        DeviceBuffer dp(count), dx(static_cast<size_t>(n) * in_features), dy(n);
        dp.Upload(params, count, st.s);
        dx.Upload(x, static_cast<size_t>(n) * in_features, st.s);
        const int rc = bm25mi_mlp_forward(dp.p, dx.p, n, in_features, dy.p, st.s);
        if (rc != BM25MI_OK)
            throw DeviceError(0, "MLP forward failed");
        dy.Download(y_out, n, st.s);
        BM25MI_MLP_HIP(hipStreamSynchronize(st.s), "MLP predict sync");
    });
}

// Minibatch Adam training on the GPU. params (in/out) start from the caller's values. The
// sample order of every epoch is a seeded shuffle; epoch_loss[e] (optional) receives the
// mean minibatch loss of epoch e. The steps of an epoch are queued without waiting on the host:
// each step's gradient partials live in this call's own buffers and the epoch's loss is summed
// on the device, read once per epoch.
extern "C" int bm25mi_mlp_train(float *params, const float *x, const float *target, int32_t n, int32_t in_features,
                                int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epoch_loss)
{
    return Invoke([&] {
        CheckShapes(x, target, n, in_features);
        Require(params && epochs >= 1 && epochs <= 100000 && batch >= 1 && batch <= n && std::isfinite(lr) && lr > 0,
                "Invalid MLP training arguments");
        const int in = in_features;
        const int64_t count = MlpParams(in);
        const int blocks = static_cast<int>(std::min<int64_t>(kBackwardGrid, (int64_t(batch) + kTile - 1) / kTile));
        Stream st;
        DeviceBuffer dp(count), dm(count), dv(count), dx(static_cast<size_t>(n) * in), dt(n), dl(1);
        DeviceBuffer partial(static_cast<size_t>(blocks) * count), lossPartial(static_cast<size_t>(blocks));
        dp.Upload(params, count, st.s);
        BM25MI_MLP_HIP(hipMemsetAsync(dm.p, 0, count * sizeof(float), st.s), "Adam m clear");
        BM25MI_MLP_HIP(hipMemsetAsync(dv.p, 0, count * sizeof(float), st.s), "Adam v clear");
        std::vector<int32_t> order(n);
        for (int32_t i = 0; i < n; ++i)
            order[i] = i;
        std::vector<float> xs(static_cast<size_t>(n) * in), ts(n);
        uint64_t rng = seed;
        int32_t step = 0;
        const int rgrid = static_cast<int>((count + 255) / 256);
        for (int e = 0; e < epochs; ++e) {
            for (int32_t i = n - 1; i > 0; --i) // Fisher-Yates
                std::swap(order[i], order[static_cast<int32_t>(SplitMix64(rng) % static_cast<uint64_t>(i + 1))]);
            BM25MI_MLP_HIP(hipStreamSynchronize(st.s), "MLP epoch sync"); // previous epoch done with xs
            for (int32_t i = 0; i < n; ++i) {
                std::memcpy(&xs[static_cast<size_t>(i) * in], &x[static_cast<size_t>(order[i]) * in], in * sizeof(float));
                ts[i] = target[order[i]];
            }
            dx.Upload(xs.data(), xs.size(), st.s);
            dt.Upload(ts.data(), ts.size(), st.s);
            BM25MI_MLP_HIP(hipMemsetAsync(dl.p, 0, sizeof(float), st.s), "MLP epoch loss clear");
            int batches = 0;
            for (int32_t b0 = 0; b0 + batch <= n; b0 += batch) {
                hipLaunchKernelGGL(KMlpBackward, dim3(blocks), dim3(kMlpThreads), 0, st.s, dp.p,
                                   dx.p + static_cast<size_t>(b0) * in, dt.p + b0, batch, in, partial.p, lossPartial.p);
                ++step;
                const float c1 = 1.0F - std::pow(0.9F, static_cast<float>(step));
                const float c2 = 1.0F - std::pow(0.999F, static_cast<float>(step));
                hipLaunchKernelGGL(KMlpReduceAdam, dim3(rgrid), dim3(256), 0, st.s, partial.p, lossPartial.p, blocks,
                                   count, batch, dp.p, dm.p, dv.p, dl.p, lr, 0.9F, 0.999F, 1e-8F, c1, c2);
                ++batches;
            }
            BM25MI_MLP_HIP(hipGetLastError(), "MLP training launches");
            float sum = 0;
            dl.Download(&sum, 1, st.s);
            BM25MI_MLP_HIP(hipStreamSynchronize(st.s), "MLP epoch loss");
            if (epoch_loss)
                epoch_loss[e] = sum / static_cast<float>(std::max(1, batches));
        }
        dp.Download(params, count, st.s);
        BM25MI_MLP_HIP(hipStreamSynchronize(st.s), "MLP train sync");
    });
}
