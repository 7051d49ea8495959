// Small kernels of the per-frame training step: the Adam update of the bandwidth network and the
// deterministic reduction of the per-tile loss partials.
#include "xdn_common.h"
#include "xdn_misc.h"

namespace xdn {

// Keras Adam (tf.keras.optimizers.Adam as used by scripts/main.py:81,303):
//   m += (g - m)(1 - b1); v += (g^2 - v)(1 - b2); p -= alpha m / (sqrt(v) + eps),
//   alpha = lr sqrt(1 - b2^t) / (1 - b1^t) (computed on the host in double).
struct AdamBatch {
    int n;
    AdamJob j[XDN_MAX_FRAMES];
    float b1, b2, eps;
};

__global__ __launch_bounds__(256) void adam_kernel(AdamBatch b)
{
    const AdamJob &J = b.j[blockIdx.y];
    const float alpha = J.alpha_dev ? *J.alpha_dev : J.alpha;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < J.count; i += gridDim.x * 256)
    {
        const float g = J.grad[i];
        float m = J.m[i], v = J.v[i];
        m += (g - m) * (1.f - b.b1);
        v += (g * g - v) * (1.f - b.b2);
        J.m[i] = m;
        J.v[i] = v;
        J.param[i] -= (m * alpha) / (sqrtf(v) + b.eps);
    }
}

struct TickBatch {
    int n;
    int *counter[XDN_MAX_FRAMES];
    float *alpha[XDN_MAX_FRAMES];
    float lr, b1, b2;
};

__global__ void adam_tick_kernel(TickBatch b)
{
    const int i = threadIdx.x;
    if (i >= b.n) return;
    const int t = *b.counter[i] + 1;
    *b.counter[i] = t;
    *b.alpha[i] = (float)((double)b.lr * sqrt(1.0 - pow((double)b.b2, (double)t)) / (1.0 - pow((double)b.b1, (double)t)));
}

hipError_t launch_adam_tick(int n, int *const *counter, float *const *alpha, float lr, float b1, float b2,
                            hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (n > XDN_MAX_FRAMES) return hipErrorInvalidValue;
    TickBatch b{};
    b.n = n;
    for (int i = 0; i < n; ++i)
    {
        b.counter[i] = counter[i];
        b.alpha[i] = alpha[i];
    }
    b.lr = lr;
    b.b1 = b1;
    b.b2 = b2;
    hipLaunchKernelGGL(adam_tick_kernel, dim3(1), dim3(64), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_adam(int n, const AdamJob *jobs, float b1, float b2, float eps, hipStream_t st)
{
    AdamBatch b{};
    b.n = n;
    int maxc = 0;
    for (int i = 0; i < n; ++i)
    {
        b.j[i] = jobs[i];
        maxc = jobs[i].count > maxc ? jobs[i].count : maxc;
    }
    b.b1 = b1;
    b.b2 = b2;
    b.eps = eps;
    if (n == 0 || maxc == 0) return hipSuccess;
    const int gx = (maxc + 255) / 256 < 256 ? (maxc + 255) / 256 : 256;
    hipLaunchKernelGGL(adam_kernel, dim3(gx, n), dim3(256), 0, st, b);
    return hipGetLastError();
}

// out[f] = scale[f] * sum(parts[f][0 .. count[f])) in a fixed order (one workgroup per frame).
struct SumBatch {
    int n;
    const float *parts[XDN_MAX_FRAMES];
    int count[XDN_MAX_FRAMES];
    float scale[XDN_MAX_FRAMES];
    float *out[XDN_MAX_FRAMES];
};

__global__ __launch_bounds__(256) void sum_kernel(SumBatch b)
{
    __shared__ float red[256];
    const int f = blockIdx.x;
    float s = 0.f;
    for (int i = threadIdx.x; i < b.count[f]; i += 256) s += b.parts[f][i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1)
    {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *b.out[f] = red[0] * b.scale[f];
}

hipError_t launch_sums(int n, const float *const *parts, const int *count, const float *scale, float *const *out,
                       hipStream_t st)
{
    SumBatch b{};
    b.n = n;
    for (int i = 0; i < n; ++i)
    {
        b.parts[i] = parts[i];
        b.count[i] = count[i];
        b.scale[i] = scale[i];
        b.out[i] = out[i];
    }
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(sum_kernel, dim3(n), dim3(256), 0, st, b);
    return hipGetLastError();
}

}  // namespace xdn
