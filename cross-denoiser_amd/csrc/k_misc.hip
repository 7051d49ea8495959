// Small kernels of the per-frame training step: the Adam update of the bandwidth network, the
// deterministic reduction of the per-tile loss partials, and the frame driver's report (the
// displayed frame and its RelL2 against the reference).
#include "xdn_common.h"
#include "xdn_misc.h"

namespace xdn {

// Keras Adam (tf.keras.optimizers.Adam as used by scripts/main.py:81,303):
//   m += (g - m)(1 - b1); v += (g^2 - v)(1 - b2); p -= alpha m / (sqrt(v) + eps),
//   alpha = lr sqrt(1 - b2^t) / (1 - b1^t) (computed in double on the host, or on the device from a
//   step counter by the loss-sum launch that precedes the update).
struct AdamBatch {
    int n;
    AdamJob j[XDN_MAX_FRAMES];
    AdamPackLayer lay[ADAM_PACK_LAYERS];
    float b1, b2, eps;
};
static_assert(sizeof(AdamBatch) <= 4096, "kernel argument too large");

typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;

// Writes the updated parameter i into its forward and data-gradient fragment slots (or the FP32
// bias block), with the same round-to-nearest-even bf16 conversion as the full packing.
__device__ __forceinline__ void pack_param(const AdamBatch &b, char *pk, int i, float x)
{
    int l = 0;
#pragma unroll
    for (int k = 1; k < ADAM_PACK_LAYERS; ++k)
        if (i >= b.lay[k].w_off) l = k;
    const AdamPackLayer &g = b.lay[l];
    if (i >= g.b_off)
    {
        reinterpret_cast<float *>(pk + g.bias_off)[i - g.b_off] = x;
        return;
    }
    const uint16_t v = (uint16_t)(__builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{x, 0.f}, bf16x2)) & 0xFFFFu);
    const int e = i - g.w_off, co = e % g.co, r = e / g.co, ci = r % g.ci, t = r / g.ci;
    {
        const int kcf = (g.ci + 15) >> 4, kc = ci >> 4, h = (ci >> 3) & 1, jj = ci & 7;
        reinterpret_cast<uint16_t *>(pk + g.fwd_off)[((t * kcf + kc) * 64 + 32 * h + co) * 8 + jj] = v;
    }
    if (g.dg_off >= 0)
    {
        const int kcd = (g.co + 15) >> 4, kc = co >> 4, h = (co >> 3) & 1, jj = co & 7;
        reinterpret_cast<uint16_t *>(pk + g.dg_off)[(((8 - t) * kcd + kc) * 64 + 32 * h + ci) * 8 + jj] = v;
    }
}

__global__ __launch_bounds__(256) void adam_kernel(AdamBatch b)
{
    const AdamJob &J = b.j[blockIdx.y];
    const float alpha = J.alpha_dev ? *J.alpha_dev : J.alpha;
    char *pk = static_cast<char *>(J.packed);
    for (int i = blockIdx.x * 256 + threadIdx.x; i < J.count; i += gridDim.x * 256)
    {
        const float g = J.grad[i];
        float m = J.m[i], v = J.v[i];
        m += (g - m) * (1.f - b.b1);
        v += (g * g - v) * (1.f - b.b2);
        J.m[i] = m;
        J.v[i] = v;
        const float p = J.param[i] - (m * alpha) / (sqrtf(v) + b.eps);
        J.param[i] = p;
        if (pk) pack_param(b, pk, i, p);
    }
}

hipError_t launch_adam(int n, const AdamJob *jobs, float b1, float b2, float eps, hipStream_t st,
                       const AdamPackLayer *layers)
{
    AdamBatch b{};
    b.n = n;
    for (int i = 0; i < n; ++i)
        if (jobs[i].packed && !layers) return hipErrorInvalidValue;
    if (layers)
        for (int l = 0; l < ADAM_PACK_LAYERS; ++l) b.lay[l] = layers[l];
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
// Optionally also the Adam step counters' advance (one extra block, one thread per counter, so its
// double-precision step size runs beside the sums).
struct SumBatch {
    int n;
    const float *parts[XDN_MAX_FRAMES];
    int count[XDN_MAX_FRAMES];
    float scale[XDN_MAX_FRAMES];
    float *out[XDN_MAX_FRAMES];
    AdamTick tick;
};
static_assert(sizeof(SumBatch) <= 4096, "kernel argument too large");

__global__ __launch_bounds__(256) void sum_kernel(SumBatch b)
{
    __shared__ float red[256];
    const int f = blockIdx.x;
    if (f == b.n)
    {
        if (threadIdx.x < b.tick.n)
        {
            const int i = threadIdx.x;
            const int t = *b.tick.counter[i] + 1;
            *b.tick.counter[i] = t;
            *b.tick.alpha[i] = (float)((double)b.tick.lr * sqrt(1.0 - pow((double)b.tick.b2, (double)t)) /
                                       (1.0 - pow((double)b.tick.b1, (double)t)));
        }
        return;
    }
    // eight loads in flight per thread; the sum keeps the order i, i + 256, ...
    float s = 0.f;
    const int cnt = b.count[f];
    const float *pp = b.parts[f];
    for (int i0 = threadIdx.x; i0 < cnt; i0 += 8 * 256)
    {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = i0 + 256 * k < cnt ? pp[i0 + 256 * k] : 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (i0 + 256 * k < cnt) s += v[k];
    }
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
                       hipStream_t st, const AdamTick *tick)
{
    SumBatch b{};
    b.n = n;
    if (tick)
    {
        if (tick->n < 0 || tick->n > XDN_MAX_FRAMES || tick->n > n) return hipErrorInvalidValue;
        b.tick = *tick;
    }
    for (int i = 0; i < n; ++i)
    {
        b.parts[i] = parts[i];
        b.count[i] = count[i];
        b.scale[i] = scale[i];
        b.out[i] = out[i];
    }
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(sum_kernel, dim3(n + (b.tick.n > 0 ? 1 : 0)), dim3(256), 0, st, b);
    return hipGetLastError();
}

// Frame report of scripts/main.py:344-346 with RelL2 of :191-196, one pass: the displayed frame
// out + emissive + envLight (written when `composed` is set), the reference ref + ref_emissive +
// ref_envLight, and per block the sum over its pixels and channels of
// (ref_c - out_c)^2 / (mean over channels of ref_c)^2 + 1e-2). Absent (null) terms are zero.
struct ErrBatch {
    int n;
    int blk_begin[XDN_MAX_FRAMES + 1];
    ErrFrame f[XDN_MAX_FRAMES];
};
static_assert(sizeof(ErrBatch) <= 4096, "kernel argument too large");

__global__ __launch_bounds__(256) void frame_error_kernel(ErrBatch b)
{
    __shared__ float red[4];
    int f = 0;
    while (f + 1 < b.n && (int)blockIdx.x >= b.blk_begin[f + 1]) ++f;
    const ErrFrame &F = b.f[f];
    const int li = blockIdx.x - b.blk_begin[f], nblk = b.blk_begin[f + 1] - b.blk_begin[f];
    const size_t np = (size_t)F.H * F.W;
    float s = 0.f;
    for (size_t p = (size_t)li * 256 + threadIdx.x; p < np; p += (size_t)nblk * 256)
    {
        float3 o = ld3(F.out, p), r = ld3(F.ref, p);
        const float *add[2] = {F.emissive, F.env_light}, *radd[2] = {F.ref_emissive, F.ref_env_light};
#pragma unroll
        for (int k = 0; k < 2; ++k)
        {
            if (add[k])
            {
                const float3 e = ld3(add[k], p);
                o.x += e.x; o.y += e.y; o.z += e.z;
            }
            if (radd[k])
            {
                const float3 e = ld3(radd[k], p);
                r.x += e.x; r.y += e.y; r.z += e.z;
            }
        }
        if (F.composed) st3(F.composed, p, o);
        const float m = (r.x + r.y + r.z) / 3.f;
        const float den = m * m + 1e-2f;
        const float dx = r.x - o.x, dy = r.y - o.y, dz = r.z - o.z;
        s += (dx * dx) / den + (dy * dy) / den + (dz * dz) / den;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) F.part[li] = (red[0] + red[1]) + (red[2] + red[3]);
}

int frame_error_blocks(int H, int W)
{
    const long np = (long)H * W;
    long nb = (np + 4095) / 4096;  // ~16 px per lane
    return (int)(nb < 1 ? 1 : (nb > 1024 ? 1024 : nb));
}

hipError_t launch_frame_error(int n, const ErrFrame *fr, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (n > XDN_MAX_FRAMES) return hipErrorInvalidValue;
    ErrBatch b{};
    b.n = n;
    int acc = 0;
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        b.blk_begin[i] = acc;
        acc += frame_error_blocks(fr[i].H, fr[i].W);
    }
    b.blk_begin[n] = acc;
    hipLaunchKernelGGL(frame_error_kernel, dim3(acc), dim3(256), 0, st, b);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void zero_kernel(float *p, size_t count)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256) p[i] = 0.f;
}

hipError_t launch_zero(float *p, size_t count, hipStream_t st)
{
    if (count == 0) return hipSuccess;
    const size_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)(blocks < 1024 ? blocks : 1024)), dim3(256), 0, st, p, count);
    return hipGetLastError();
}

}  // namespace xdn
