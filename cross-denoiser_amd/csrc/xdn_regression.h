// Device helpers shared by the cross-regression forward and backward kernels: the 6x6 ridge
// systems, their Cholesky solve and the regression weight (regression.cu.cc:197-287).
#pragma once
#include "xdn_common.h"

namespace xdn {

// Upper-triangular 6x6 normal equations of one channel plus the right-hand side.
struct Sys6 {
    float a[21];  // row-major upper triangle: (0,0) (0,1) .. (0,5) (1,1) .. (5,5)
    float x[6];
};

__device__ __forceinline__ constexpr int up6(int r, int c) { return r * 6 - (r * (r - 1)) / 2 + (c - r); }

// Ridge + Cholesky (cholesky_color :206-218 with the SQRT clamp :197-204) + forward/backward
// substitution (:220-242) for one channel; all indices compile-time.
__device__ __forceinline__ void solve6(const Sys6 &s, float beta[6])
{
    float L[6][6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j)
        {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < j; ++k) acc += L[i][k] * L[j][k];
            const float aji = s.a[up6(j, i)] + (i == j ? 1e-3f : 0.f);
            if (i == j)
                L[i][i] = sqrtf(fmaxf(1e-5f, aji - acc));
            else
                L[i][j] = (1.f / L[j][j]) * (aji - acc);
        }
    float w[6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
    {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < i; ++k) acc += L[i][k] * w[k];
        w[i] = (s.x[i] - acc) / L[i][i];
    }
#pragma unroll
    for (int i = 5; i >= 0; --i)
    {
        float acc = 0.f;
#pragma unroll
        for (int k = 5; k > i; --k) acc += L[k][i] * beta[k];
        beta[i] = (w[i] - acc) / L[i][i];
    }
}

// The 61 distinct accumulators of the three per-channel 6x6 systems. Feature order per channel:
// [1, d1_c, d2_c, n0, n1, n2]; entries among {0,3,4,5} do not depend on the channel.
struct RegAcc {
    float s00, s03, s04, s05, s33, s34, s35, s44, s45, s55;  // shared
    float c01[3], c02[3], c11[3], c12[3], c22[3], c13[3], c14[3], c15[3], c23[3], c24[3], c25[3];
    float x0[3], x1[3], x2[3], x3[3], x4[3], x5[3];

    __device__ __forceinline__ void zero()
    {
        s00 = s03 = s04 = s05 = s33 = s34 = s35 = s44 = s45 = s55 = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            c01[c] = c02[c] = c11[c] = c12[c] = c22[c] = c13[c] = c14[c] = c15[c] = c23[c] = c24[c] =
                c25[c] = x0[c] = x1[c] = x2[c] = x3[c] = x4[c] = x5[c] = 0.f;
    }

    // One tap: weight w, 1/(std_c + std_i + 1e-2), own-half colour difference dw (tap - center),
    // albedo difference d2, normal difference n, target colour t.
    __device__ __forceinline__ void add(float w, float inv, float3 dw, float3 d2, float3 n, float3 t)
    {
        const float wn0 = w * n.x, wn1 = w * n.y, wn2 = w * n.z;
        s00 += w;
        s03 += wn0;
        s04 += wn1;
        s05 += wn2;
        s33 += wn0 * n.x;
        s34 += wn0 * n.y;
        s35 += wn0 * n.z;
        s44 += wn1 * n.y;
        s45 += wn1 * n.z;
        s55 += wn2 * n.z;
        const float d1v[3] = {dw.x * inv, dw.y * inv, dw.z * inv};
        const float d2v[3] = {d2.x, d2.y, d2.z};
        const float tv[3] = {t.x, t.y, t.z};
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
            const float d1 = d1v[c], dd2 = d2v[c], tt = tv[c];
            const float wd1 = w * d1, wd2 = w * dd2;
            c01[c] += wd1;
            c02[c] += wd2;
            c11[c] += wd1 * d1;
            c12[c] += wd1 * dd2;
            c22[c] += wd2 * dd2;
            c13[c] += wd1 * n.x;
            c14[c] += wd1 * n.y;
            c15[c] += wd1 * n.z;
            c23[c] += wd2 * n.x;
            c24[c] += wd2 * n.y;
            c25[c] += wd2 * n.z;
            x0[c] += w * tt;
            x1[c] += wd1 * tt;
            x2[c] += wd2 * tt;
            x3[c] += wn0 * tt;
            x4[c] += wn1 * tt;
            x5[c] += wn2 * tt;
        }
    }

    // System of channel c (c compile-time).
    __device__ __forceinline__ void system(int c, Sys6 &s) const
    {
        // row 0
        s.a[up6(0, 0)] = s00; s.a[up6(0, 1)] = c01[c]; s.a[up6(0, 2)] = c02[c];
        s.a[up6(0, 3)] = s03; s.a[up6(0, 4)] = s04; s.a[up6(0, 5)] = s05;
        s.a[up6(1, 1)] = c11[c]; s.a[up6(1, 2)] = c12[c]; s.a[up6(1, 3)] = c13[c];
        s.a[up6(1, 4)] = c14[c]; s.a[up6(1, 5)] = c15[c];
        s.a[up6(2, 2)] = c22[c]; s.a[up6(2, 3)] = c23[c]; s.a[up6(2, 4)] = c24[c]; s.a[up6(2, 5)] = c25[c];
        s.a[up6(3, 3)] = s33; s.a[up6(3, 4)] = s34; s.a[up6(3, 5)] = s35;
        s.a[up6(4, 4)] = s44; s.a[up6(4, 5)] = s45;
        s.a[up6(5, 5)] = s55;
        s.x[0] = x0[c]; s.x[1] = x1[c]; s.x[2] = x2[c]; s.x[3] = x3[c]; s.x[4] = x4[c]; s.x[5] = x5[c];
    }
};

// Regression weight (regressionWgt :278-287, +1e-7 at :432) with the +1e-4 folded into cvar_e.
__device__ __forceinline__ float reg_weight(float3 dw, float cvar_e, float ivar)
{
    const float n2 = dw.x * dw.x + dw.y * dw.y + dw.z * dw.z;
    return fast_exp2(-kLog2e * n2 * fast_rcp(cvar_e + ivar)) + 1e-7f;
}

}  // namespace xdn
