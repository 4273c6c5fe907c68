// Training and inference kernels of the learned K-th-score model (no reference
// counterpart): an MLP in -> 64 -> 64 -> 1 with ReLU, FP32 throughout.
//
// Parameter layout (flat): W1[64][in] b1[64] W2[64][64] b2[64] W3[64] b3[1].
//   forward   one thread per sample, weights staged in LDS (broadcast reads).
//   backward  each workgroup walks 64-sample tiles: tile forward pass into LDS,
//             mean-squared-error gradient, and per-thread register accumulators
//             for its share of every parameter gradient; partials go to a
//             per-workgroup buffer and a second kernel sums them in a fixed
//             order, so gradients are deterministic run to run.
//   adam      elementwise Adam update with bias correction.
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

// Partial gradients of one workgroup: partial[blockIdx.x][0, P) and lossPartial[blockIdx.x].
__global__ __launch_bounds__(kMlpThreads) void KMlpBackward(const float *__restrict__ p, const float *__restrict__ x,
                                                            const float *__restrict__ target, int n, int in,
                                                            float *__restrict__ partial,
                                                            float *__restrict__ lossPartial)
{
    __shared__ float sX[kTile * kMaxIn];
    __shared__ float sH1[kTile * kH]; // h1, later dz1
    __shared__ float sH2[kTile * kH]; // h2, later dz2
    __shared__ float sW2[kH * kH];
    __shared__ float sDy[kTile];
    __shared__ float sRed[kMlpThreads / 64];
    const float *W1 = p, *B1 = W1 + kH * in, *W2 = B1 + kH, *B2 = W2 + kH * kH, *W3 = B2 + kH, *B3 = W3 + kH;
    const int tid = threadIdx.x;
    for (int i = tid; i < kH * kH; i += kMlpThreads)
        sW2[i] = W2[i];
    // Ownership: unit j = tid % 64 and rows r = tid / 64 + 4 * m (m < 16) of every 64x64 tile.
    const int unit = tid & 63, quarter = tid >> 6;
    float gW2[16];   // dW2[unit][quarter*16 + m]
    float gW1[kMaxIn / 4 + 1]; // dW1[unit][quarter + 4*m], in/4 rounded up
    float gB1 = 0, gB2 = 0, gW3 = 0, gB3 = 0, loss = 0;
#pragma unroll
    for (int m = 0; m < 16; ++m)
        gW2[m] = 0;
#pragma unroll
    for (int m = 0; m < kMaxIn / 4 + 1; ++m)
        gW1[m] = 0;
    const float scale = 2.0F / static_cast<float>(n);
    __syncthreads();
    for (int64_t t0 = int64_t(blockIdx.x) * kTile; t0 < n; t0 += int64_t(gridDim.x) * kTile) {
        const int rows = static_cast<int>(min<int64_t>(kTile, n - t0));
        for (int i = tid; i < kTile * in; i += kMlpThreads) {
            const int s = i / in;
            sX[i] = s < rows ? x[(t0 + s) * in + (i - s * in)] : 0.0F;
        }
        __syncthreads();
        // h1 = relu(W1 x + b1)
        for (int s = quarter; s < kTile; s += 4) {
            float z = B1[unit];
            for (int i = 0; i < in; ++i)
                z = fmaf(W1[unit * in + i], sX[s * in + i], z);
            sH1[s * kH + unit] = fmaxf(z, 0.0F);
        }
        __syncthreads();
        // h2 = relu(W2 h1 + b2)
        for (int s = quarter; s < kTile; s += 4) {
            float z = B2[unit];
            for (int j = 0; j < kH; ++j)
                z = fmaf(sW2[unit * kH + j], sH1[s * kH + j], z);
            sH2[s * kH + unit] = fmaxf(z, 0.0F);
        }
        __syncthreads();
        // y, dy = 2 (y - target) / n
        if (tid < kTile) {
            float yv = B3[0];
            for (int k = 0; k < kH; ++k)
                yv = fmaf(W3[k], sH2[tid * kH + k], yv);
            const float d = tid < rows ? yv - target[t0 + tid] : 0.0F;
            loss = fmaf(d, d, loss);
            sDy[tid] = d * scale;
        }
        __syncthreads();
        // dW3[unit], db3, then dz2 = dy * W3 * (h2 > 0) in place of h2.
        if (quarter == 0) {
            for (int s = 0; s < kTile; ++s)
                gW3 = fmaf(sDy[s], sH2[s * kH + unit], gW3);
            if (unit == 0)
                for (int s = 0; s < kTile; ++s)
                    gB3 += sDy[s];
        }
        __syncthreads();
        const float w3 = W3[unit];
        for (int s = quarter; s < kTile; s += 4)
            sH2[s * kH + unit] = sH2[s * kH + unit] > 0.0F ? sDy[s] * w3 : 0.0F;
        __syncthreads();
        // dW2[k][j] += sum_s dz2[s][k] h1[s][j]; owner k = unit, j = quarter*16 + m.
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int j = quarter * 16 + m;
            float acc = gW2[m];
            for (int s = 0; s < kTile; ++s)
                acc = fmaf(sH2[s * kH + unit], sH1[s * kH + j], acc);
            gW2[m] = acc;
        }
        if (quarter == 0)
            for (int s = 0; s < kTile; ++s)
                gB2 += sH2[s * kH + unit];
        __syncthreads();
        // dz1[s][j] = (h1 > 0) * sum_k dz2[s][k] W2[k][j], in place of h1 (same owner reads then writes).
        for (int s = quarter; s < kTile; s += 4) {
            float acc = 0.0F;
            for (int k = 0; k < kH; ++k)
                acc = fmaf(sH2[s * kH + k], sW2[k * kH + unit], acc);
            sH1[s * kH + unit] = sH1[s * kH + unit] > 0.0F ? acc : 0.0F;
        }
        __syncthreads();
        // dW1[j][i] += sum_s dz1[s][j] x[s][i]; owner j = unit, i = quarter + 4*m.
        for (int m = 0; quarter + 4 * m < in; ++m) {
            const int i = quarter + 4 * m;
            float acc = gW1[m];
            for (int s = 0; s < kTile; ++s)
                acc = fmaf(sH1[s * kH + unit], sX[s * in + i], acc);
            gW1[m] = acc;
        }
        if (quarter == 0)
            for (int s = 0; s < kTile; ++s)
                gB1 += sH1[s * kH + unit];
        __syncthreads();
    }
    float *out = partial + int64_t(blockIdx.x) * MlpParams(in);
    float *oW1 = out, *oB1 = oW1 + kH * in, *oW2 = oB1 + kH, *oB2 = oW2 + kH * kH, *oW3 = oB2 + kH, *oB3 = oW3 + kH;
#pragma unroll
    for (int m = 0; m < 16; ++m)
        oW2[unit * kH + quarter * 16 + m] = gW2[m];
    for (int m = 0; quarter + 4 * m < in; ++m)
        oW1[unit * in + quarter + 4 * m] = gW1[m];
    if (quarter == 0) {
        oB1[unit] = gB1;
        oB2[unit] = gB2;
        oW3[unit] = gW3;
        if (unit == 0)
            oB3[0] = gB3;
    }
    // Loss partial: fixed-order block sum.
    for (int o = 32; o > 0; o >>= 1)
        loss += __shfl_xor(loss, o);
    if ((tid & 63) == 0)
        sRed[tid >> 6] = loss;
    __syncthreads();
    if (tid == 0)
        lossPartial[blockIdx.x] = sRed[0] + sRed[1] + sRed[2] + sRed[3];
}

// Sums the workgroup partials in a fixed order. train = 0: grad += sum, lossOut[0] = mean loss;
// train = 1 (training steps): grad = sum, lossOut[0] += mean loss (the epoch's running sum).
__global__ void KMlpReduce(const float *__restrict__ partial, const float *__restrict__ lossPartial, int blocks,
                           int64_t count, int n, float *__restrict__ grad, float *__restrict__ lossOut, int train)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) {
        float s = 0.0F;
        for (int b = 0; b < blocks; ++b)
            s += partial[int64_t(b) * count + i];
        grad[i] = train ? s : grad[i] + s;
    }
    if (i == 0 && lossOut) {
        float s = 0.0F;
        for (int b = 0; b < blocks; ++b)
            s += lossPartial[b];
        const float mean = s / static_cast<float>(n);
        lossOut[0] = train ? lossOut[0] + mean : mean;
    }
}

__global__ void KMlpAdam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                         float *__restrict__ v, int64_t count, float lr, float beta1, float beta2, float eps,
                         float c1, float c2)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    const float gi = g[i];
    const float mi = beta1 * m[i] + (1.0F - beta1) * gi;
    const float vi = beta2 * v[i] + (1.0F - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

namespace {
std::mutex scratchMutex;
float *scratch = nullptr;
size_t scratchFloats = 0;
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
        const size_t need = static_cast<size_t>(blocks) * (count + 1);
        if (need > scratchFloats) {
            auto s = static_cast<hipStream_t>(stream);
            BM25MI_MLP_HIP(hipStreamSynchronize(s), "MLP scratch sync");
            if (scratch)
                BM25MI_MLP_HIP(hipFree(scratch), "MLP scratch free");
            scratch = nullptr;
            scratchFloats = 0;
            BM25MI_MLP_HIP(hipMalloc(&scratch, need * sizeof(float)), "MLP scratch alloc");
            scratchFloats = need;
        }
        float *partial = scratch, *lossPartial = scratch + static_cast<size_t>(blocks) * count;
        auto s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(KMlpBackward, dim3(blocks), dim3(kMlpThreads), 0, s, params, x, target, n, in_features,
                           partial, lossPartial);
        BM25MI_MLP_HIP(hipGetLastError(), "MLP backward launch");
        const int rgrid = static_cast<int>((count + 255) / 256);
        hipLaunchKernelGGL(KMlpReduce, dim3(rgrid), dim3(256), 0, s, partial, lossPartial, blocks, count, n, grad,
                           loss_out, 0);
        BM25MI_MLP_HIP(hipGetLastError(), "MLP reduce launch");
        // The scratch buffer is reused by the next call: finish this one first.
        BM25MI_MLP_HIP(hipStreamSynchronize(s), "MLP backward sync");
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
        DeviceBuffer dp(count), dg(count), dm(count), dv(count), dx(static_cast<size_t>(n) * in), dt(n), dl(1);
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
        const int agrid = rgrid;
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
                hipLaunchKernelGGL(KMlpReduce, dim3(rgrid), dim3(256), 0, st.s, partial.p, lossPartial.p, blocks,
                                   count, batch, dg.p, dl.p, 1);
                ++step;
                const float c1 = 1.0F - std::pow(0.9F, static_cast<float>(step));
                const float c2 = 1.0F - std::pow(0.999F, static_cast<float>(step));
                hipLaunchKernelGGL(KMlpAdam, dim3(agrid), dim3(256), 0, st.s, dp.p, dg.p, dm.p, dv.p, count, lr, 0.9F,
                                   0.999F, 1e-8F, c1, c2);
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
