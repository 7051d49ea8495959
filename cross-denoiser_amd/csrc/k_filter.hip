// Bandwidth-driven bilateral spatial filter on gfx950 (SpatialFilterKernel / SpatialFilterGradKernel,
// regression.cu.cc:33-195) and the fused per-frame training kernel.
//
// Weights w^X = exp(-b0|dX|^2 - b2|dalb|^2 - b3|dn|^2 - b4 ((dx/h)^2 + (dy/h)^2)) over an 11x11
// reflect-101 window, X in {A (band 0), B (band 1)}; the bands are pre-scaled by -log2(e) so each
// weight is one v_exp_f32. Nothing per tap is stored: every backward recomputes the weights.
//
// Kernels (16x16 output tile per 256-lane workgroup, the tile plus a 5-px halo staged in LDS as
// three float4 planes per pixel: (A.xyz, B.x), (B.yz, alb.xy), (alb.z, n.xyz)):
//   sf_tile<FWD>     unnormalized sums and weight sums                          VALU-bound
//   sf_tile<GRAD>    band gradient from upstream grads, weights recomputed     VALU-bound
//   sf_tile<TRAIN>   forward sums + the 32 band-gradient moments in registers, then the temporal
//                    accumulation, loss and the whole pointwise backward (aggregation ->
//                    weights -> band transform) in the epilogue; writes the output image and
//                    d loss / d network output                                  VALU-bound
//   sf_grad_inputs   gradient w.r.t. the four images, gather form (no atomics): every pixel
//                    collects its own-center term and the terms of all centers whose (mirrored)
//                    window contains it
//   sf_generic       any window size, direct loads (forward / grad)
#include "xdn_common.h"
#include "xdn_kernels.h"

namespace xdn {

enum { SF_FWD = 0, SF_GRAD = 1, SF_TRAIN = 2 };

struct SfBatch {
    TileMap map;
    SfFrame f[XDN_MAX_FRAMES];
    int half;
};
struct TrBatch {
    TileMap map;
    TrainFrame f[XDN_MAX_FRAMES];
};
static_assert(sizeof(SfBatch) <= 4096, "kernel argument too large");
static_assert(sizeof(TrBatch) <= 4096, "kernel argument too large");

constexpr int SF_T = 16;

// Per-tap state of one center pixel.
struct SfCenter {
    float3 cA, cB, ca, cn;
    float k0, k1, k2, k3, k4;  // -log2e * band, k4 already divided by half^2
};

struct SfAcc {
    float3 aA, aB;
    float sA, sB;
    // GRAD: band-gradient accumulators; TRAIN: moments sum w*d_k*(c,1)
    float g[5];
    float4 MA[4], MB[4];
};

template <int MODE>
__device__ __forceinline__ void sf_tap(const SfCenter &C, SfAcc &acc, float3 iA, float3 iB, float3 ia, float3 in,
                                       float pd, float kpd, float3 gA, float hA, float3 gB, float hB)
{
    const float dA = dist2(C.cA, iA), dB = dist2(C.cB, iB), dal = dist2(C.ca, ia), dn = dist2(C.cn, in);
    const float sh = C.k2 * dal + (C.k3 * dn + kpd);
    const float wA = fast_exp2(C.k0 * dA + sh);
    const float wB = fast_exp2(C.k1 * dB + sh);
    if (MODE == SF_FWD || MODE == SF_TRAIN)
    {
        acc.aA.x += wA * iA.x; acc.aA.y += wA * iA.y; acc.aA.z += wA * iA.z;
        acc.aB.x += wB * iB.x; acc.aB.y += wB * iB.y; acc.aB.z += wB * iB.z;
        acc.sA += wA;
        acc.sB += wB;
    }
    if (MODE == SF_GRAD)
    {
        const float uA = wA * (gA.x * iA.x + gA.y * iA.y + gA.z * iA.z + hA);
        const float uB = wB * (gB.x * iB.x + gB.y * iB.y + gB.z * iB.z + hB);
        const float uS = uA + uB;
        acc.g[0] += uA * dA;
        acc.g[1] += uB * dB;
        acc.g[2] += uS * dal;
        acc.g[3] += uS * dn;
        acc.g[4] += uS * pd;
    }
    if (MODE == SF_TRAIN)
    {
        const float mA[4] = {wA * dA, wA * dal, wA * dn, wA * pd};
        const float mB[4] = {wB * dB, wB * dal, wB * dn, wB * pd};
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            acc.MA[k].x += mA[k] * iA.x; acc.MA[k].y += mA[k] * iA.y; acc.MA[k].z += mA[k] * iA.z;
            acc.MA[k].w += mA[k];
            acc.MB[k].x += mB[k] * iB.x; acc.MB[k].y += mB[k] * iB.y; acc.MB[k].z += mB[k] * iB.z;
            acc.MB[k].w += mB[k];
        }
    }
}

__device__ __forceinline__ void sf_zero(SfAcc &a)
{
    a.aA = a.aB = make_float3(0.f, 0.f, 0.f);
    a.sA = a.sB = 0.f;
#pragma unroll
    for (int k = 0; k < 5; ++k) a.g[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) a.MA[k] = a.MB[k] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__device__ __forceinline__ void sf_center_bands(SfCenter &C, const float *band, size_t c, int half)
{
    const float *b = band + 5 * c;
    C.k0 = -kLog2e * b[0];
    C.k1 = -kLog2e * b[1];
    C.k2 = -kLog2e * b[2];
    C.k3 = -kLog2e * b[3];
    C.k4 = -kLog2e * b[4] / (float)(half * half);
}

// Epilogue of the fused training kernel for one pixel; returns the pixel's loss sum.
// Follows aggregation_loss (scripts/main.py:214-248), its analytic gradient, the band-gradient
// contraction (regression.cu.cc:58-80) and the band/alpha transform (main.py:111-113).
__device__ __forceinline__ float train_epilogue(const TrainFrame &F, uint32_t c, const SfAcc &acc, float inv_count,
                                                float3 lab2, float3 lab1)
{
    // the five bandwidth outputs are read again at the end, where the band transform needs them:
    // kept live from here they push the epilogue past the registers of five resident blocks
    const float *pr = F.param + 6 * (size_t)c;
    const float alpha = 1.f / (1.f + __expf(-pr[5]));
    const float d1 = acc.sA + 1e-4f, d2 = acc.sB + 1e-4f;
    const float i1 = 1.f / d1, i2 = 1.f / d2, i12 = 1.f / (acc.sA + acc.sB + 1e-4f);
    const float op = F.opacity[c] > 0.f ? 1.f : 0.f;
    const float s = F.succ_e[c] * F.succ_prev[c];
    const float a = s * alpha + (1.f - s);
    const float3 pv = ld3(F.prev_out, c);
    const float3 plab1 = ld3(F.prev_ols2, c), plab2 = ld3(F.prev_ols1, c);
    const float o1v[3] = {acc.aA.x * i1, acc.aA.y * i1, acc.aA.z * i1};
    const float o2v[3] = {acc.aB.x * i2, acc.aB.y * i2, acc.aB.z * i2};
    const float ov[3] = {(acc.aA.x + acc.aB.x) * i12, (acc.aA.y + acc.aB.y) * i12, (acc.aA.z + acc.aB.z) * i12};
    const float pvv[3] = {pv.x, pv.y, pv.z};
    const float l1v[3] = {lab1.x, lab1.y, lab1.z}, l2v[3] = {lab2.x, lab2.y, lab2.z};
    const float p1v[3] = {plab1.x, plab1.y, plab1.z}, p2v[3] = {plab2.x, plab2.y, plab2.z};
    const float cg = 0.25f * op * inv_count;
    float lsum = 0.f, ga = 0.f;
    float outl[3], outv[3], gA[3], gB[3], lossv[3];
    float hA = 0.f, hB = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        const float o = (a * ov[k] + (1.f - a) * pvv[k]) * op;
        outl[k] = o;
        outv[k] = fmaxf(0.f, expm1f(o));
        const float o1 = a * o1v[k] + (1.f - a) * pvv[k];
        const float o2 = a * o2v[k] + (1.f - a) * pvv[k];
        const float r1 = 1.f / (l1v[k] * l1v[k] + 1e-2f), r2 = 1.f / (l2v[k] * l2v[k] + 1e-2f);
        const float q1 = 1.f / (p1v[k] * p1v[k] + 1e-2f), q2 = 1.f / (p2v[k] * p2v[k] + 1e-2f);
        const float e1 = o1 - l1v[k], e2 = o2 - l2v[k], f1 = o1 - p1v[k], f2 = o2 - p2v[k];
        const float l = ((e1 * e1 * r1 + e2 * e2 * r2) * 0.5f + (f1 * f1 * q1 + f2 * f2 * q2) * 0.5f) * 0.5f * op;
        lossv[k] = l;
        lsum += l;
        const float g1 = cg * 2.f * (e1 * r1 + f1 * q1);
        const float g2 = cg * 2.f * (e2 * r2 + f2 * q2);
        ga += g1 * (o1v[k] - pvv[k]) + g2 * (o2v[k] - pvv[k]);
        const float go1 = a * g1, go2 = a * g2;
        gA[k] = go1 * i1;
        gB[k] = go2 * i2;
        hA -= go1 * o1v[k] * i1;  // d out1 / d sumw1 = -sum1 / d1^2 = -out1 / d1
        hB -= go2 * o2v[k] * i2;
    }
    st3(F.out_log, c, make_float3(outl[0], outl[1], outl[2]));
    st3(F.out, c, make_float3(outv[0], outv[1], outv[2]));
    if (F.loss) st3(F.loss, c, make_float3(lossv[0], lossv[1], lossv[2]));
    // band gradients: d w / d b_k = -d_k w  ->  dL/db_k = -(<g, M_k.xyz> + h M_k.w)
    float oA[4], oB[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        oA[k] = -(gA[0] * acc.MA[k].x + gA[1] * acc.MA[k].y + gA[2] * acc.MA[k].z + hA * acc.MA[k].w);
        oB[k] = -(gB[0] * acc.MB[k].x + gB[1] * acc.MB[k].y + gB[2] * acc.MB[k].z + hB * acc.MB[k].w);
    }
    const float gb[5] = {oA[0], oB[0], oA[1] + oB[1], oA[2] + oB[2], oA[3] + oB[3]};
    float x[5];  // all five loaded before the first store (the stores may alias them as far as the
#pragma unroll  // compiler knows)
    for (int k = 0; k < 5; ++k) x[k] = pr[k];
    float *dp = F.dparam + 6 * (size_t)c;
#pragma unroll
    for (int k = 0; k < 5; ++k)
    {
        const float band = 1.f / (x[k] * x[k] + 1e-4f);
        dp[k] = gb[k] * (-2.f * x[k] * band * band);
    }
    dp[5] = s * ga * alpha * (1.f - alpha);
    return lsum;
}

template <int HALF, int MODE, class Batch>
__device__ __forceinline__ void sf_tile_body(const Batch &b, float *red)
{
    constexpr int E = SF_T + 2 * HALF;
    __shared__ float4 s0[E * E], s1[E * E], s2[E * E];
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const auto &F = b.f[L.f];
    const int H = F.H, W = F.W;
    const int x0 = L.tx * SF_T, y0 = L.ty * SF_T;
    const float *IA, *IB;
    if constexpr (MODE == SF_TRAIN) { IA = F.ols1; IB = F.ols2; }
    else { IA = F.ia; IB = F.ib; }
    for (int i = threadIdx.x; i < E * E; i += 256)
    {
        const int ly = i / E, lx = i - ly * E;
        // halo rows/columns past the last live pixel + HALF are never read: clamp them so the mirror
        // stays inside the image for partial tiles
        const int gy = mirror(min(y0 + ly - HALF, H - 1 + HALF), H), gx = mirror(min(x0 + lx - HALF, W - 1 + HALF), W);
        const size_t p = (size_t)gy * W + gx;
        const float3 a = ld3(IA, p), bb = ld3(IB, p), al = ld3(F.alb, p), n = ld3(F.nrm, p);
        s0[i] = make_float4(a.x, a.y, a.z, bb.x);
        s1[i] = make_float4(bb.y, bb.z, al.x, al.y);
        s2[i] = make_float4(al.z, n.x, n.y, n.z);
    }
    __syncthreads();
    const int lx = threadIdx.x & (SF_T - 1), ly = threadIdx.x / SF_T;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cx = x0 + lx, cy = y0 + ly;
    const bool live = cx < W && cy < H;
    float lsum = 0.f;
    if (live)
    {
        const uint32_t c = (uint32_t)cy * (uint32_t)W + (uint32_t)cx;  // a frame has < 2^32 pixels
        const int q0 = (ly + HALF) * E + lx + HALF;
        SfCenter C;
        {
            const float4 P0 = s0[q0], P1 = s1[q0], P2 = s2[q0];
            C.cA = make_float3(P0.x, P0.y, P0.z);
            C.cB = make_float3(P0.w, P1.x, P1.y);
            C.ca = make_float3(P1.z, P1.w, P2.x);
            C.cn = make_float3(P2.y, P2.z, P2.w);
        }
        if constexpr (MODE == SF_TRAIN)
        {
            const float *pr = F.param + 6 * (size_t)c;
            float bnd[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) bnd[k] = 1.f / (pr[k] * pr[k] + 1e-4f);
            C.k0 = -kLog2e * bnd[0]; C.k1 = -kLog2e * bnd[1]; C.k2 = -kLog2e * bnd[2];
            C.k3 = -kLog2e * bnd[3]; C.k4 = -kLog2e * bnd[4] / (float)(HALF * HALF);
        }
        else
            sf_center_bands(C, F.band, c, HALF);
        float3 gA = make_float3(0.f, 0.f, 0.f), gB = gA;
        float hA = 0.f, hB = 0.f;
        if constexpr (MODE == SF_GRAD)
        {
            gA = ld3(F.gaccA, c); gB = ld3(F.gaccB, c); hA = F.gwA[c]; hB = F.gwB[c];
        }
        SfAcc acc;
        sf_zero(acc);
        constexpr float inv_h2 = 1.f / (float)(HALF * HALF);
        for (int dy = -HALF; dy <= HALF; ++dy)
        {
            const float pdy = (float)(dy * dy) * inv_h2;
            const float kdy = C.k4 * (float)(dy * dy);
            const int base = (ly + HALF + dy) * E + lx + HALF;
#pragma unroll
            for (int dx = -HALF; dx <= HALF; ++dx)
            {
                const int q = base + dx;
                const float4 P0 = s0[q], P1 = s1[q], P2 = s2[q];
                const float pd = pdy + (float)(dx * dx) * inv_h2;
                const float kpd = kdy + C.k4 * (float)(dx * dx);
                sf_tap<MODE>(C, acc, make_float3(P0.x, P0.y, P0.z), make_float3(P0.w, P1.x, P1.y),
                             make_float3(P1.z, P1.w, P2.x), make_float3(P2.y, P2.z, P2.w), pd, kpd, gA, hA, gB, hB);
            }
        }
        if constexpr (MODE == SF_FWD)
        {
            st3(F.accA, c, acc.aA);
            st3(F.accB, c, acc.aB);
            F.wsA[c] = acc.sA;
            F.wsB[c] = acc.sB;
        }
        else if constexpr (MODE == SF_GRAD)
        {
            float *o = F.gband + 5 * (size_t)c;
#pragma unroll
            for (int k = 0; k < 5; ++k) o[k] = -acc.g[k];
        }
        else
        {
            const float inv_count = 1.f / (3.f * (float)H * (float)W);
            lsum = train_epilogue(F, c, acc, inv_count, C.cA, C.cB);
        }
    }
    if constexpr (MODE == SF_TRAIN)
    {
        // deterministic per-tile loss partial: wave reduction then 4 waves through LDS (wave and lane
        // from a wave-uniform value and the lane count, so no thread index stays live across the taps)
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off);
        const bool lane0 = __lane_id() == 0;
        if (lane0) red[wave] = lsum;
        __syncthreads();
        if (wave == 0 && lane0) F.loss_part[tile - b.map.begin[L.f]] = (red[0] + red[1]) + (red[2] + red[3]);
    }
}

template <int HALF, int MODE>
__global__ __launch_bounds__(256) void sf_tile(SfBatch b)
{
    sf_tile_body<HALF, MODE>(b, nullptr);
}

// Five resident blocks per CU (32.5 KB of LDS each): 96 VGPRs instead of 106. The epilogue is kept
// lean for it (see train_epilogue); one 8-byte value is still spilled there, none in the tap loop.
template <int HALF>
__global__ __launch_bounds__(256, 5) void sf_train_tile(TrBatch b)
{
    __shared__ float red[4];
    sf_tile_body<HALF, SF_TRAIN>(b, red);
}

// Any window size: direct loads with the mirror boundary.
template <int MODE>
__global__ __launch_bounds__(256) void sf_generic(SfBatch b)
{
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const SfFrame &F = b.f[L.f];
    const int H = F.H, W = F.W, HALF = b.half;
    const int cx = L.tx * SF_T + (threadIdx.x & 15), cy = L.ty * SF_T + (threadIdx.x >> 4);
    if (cx >= W || cy >= H) return;
    const size_t c = (size_t)cy * W + cx;
    SfCenter C;
    C.cA = ld3(F.ia, c); C.cB = ld3(F.ib, c); C.ca = ld3(F.alb, c); C.cn = ld3(F.nrm, c);
    sf_center_bands(C, F.band, c, HALF);
    float3 gA = make_float3(0.f, 0.f, 0.f), gB = gA;
    float hA = 0.f, hB = 0.f;
    if (MODE == SF_GRAD)
    {
        gA = ld3(F.gaccA, c); gB = ld3(F.gaccB, c); hA = F.gwA[c]; hB = F.gwB[c];
    }
    SfAcc acc;
    sf_zero(acc);
    const float inv_h2 = 1.f / (float)(HALF * HALF);
    for (int dy = -HALF; dy <= HALF; ++dy)
        for (int dx = -HALF; dx <= HALF; ++dx)
        {
            const size_t i = (size_t)mirror(cy + dy, H) * W + mirror(cx + dx, W);
            const float r2 = (float)(dx * dx + dy * dy);
            sf_tap<MODE>(C, acc, ld3(F.ia, i), ld3(F.ib, i), ld3(F.alb, i), ld3(F.nrm, i), r2 * inv_h2,
                         C.k4 * r2, gA, hA, gB, hB);
        }
    if (MODE == SF_FWD)
    {
        st3(F.accA, c, acc.aA);
        st3(F.accB, c, acc.aB);
        F.wsA[c] = acc.sA;
        F.wsB[c] = acc.sB;
    }
    else
    {
        float *o = F.gband + 5 * c;
#pragma unroll
        for (int k = 0; k < 5; ++k) o[k] = -acc.g[k];
    }
}

// ---------------------------------------------------------------------------------------------
// Gradient w.r.t. the images (the reference returns None for them, scripts/ops.py:32).
// For every pixel q:
//   own term   (q as the center): for each tap i = mirror(q+d), with u = w (<g_q, c_i> + h_q):
//              dq += -2 b_q u (c_q - c_i)  (per image and band)
//   tap terms  (q as the tap of center p, mirror(p+d) = q):
//              dq += w g_p (direct, colour images only) + 2 b_p u_pq (c_p - c_q)
// The centers p whose mirrored window contains q are enumerated per dimension: p = x - d for
// x in {q, -q (q > 0), 2n-2-q (q <= n-2)} with 0 <= p < n.
// LDS: the tile plus halo of every per-pixel quantity (25 floats), 26x26 pixels for win 11.
// ---------------------------------------------------------------------------------------------
template <int HALF>
__global__ __launch_bounds__(256) void sf_grad_inputs_tile(SfBatch b)
{
    constexpr int E = SF_T + 2 * HALF;
    constexpr int NQ = 7;  // float4 per pixel
    // per pixel: (A.xyz, B.x) (B.yz, alb.xy) (alb.z, n.xyz) (b0..b3) (b4, gA.xyz) (hA, gB.xyz) (hB, -, -, -)
    __shared__ float4 s[NQ][E * E];
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const SfFrame &F = b.f[L.f];
    const int H = F.H, W = F.W;
    const int x0 = L.tx * SF_T, y0 = L.ty * SF_T;
    for (int i = threadIdx.x; i < E * E; i += 256)
    {
        const int ly = i / E, lx = i - ly * E;
        const int gy = mirror(min(y0 + ly - HALF, H - 1 + HALF), H), gx = mirror(min(x0 + lx - HALF, W - 1 + HALF), W);
        const size_t p = (size_t)gy * W + gx;
        const float3 a = ld3(F.ia, p), bb = ld3(F.ib, p), al = ld3(F.alb, p), n = ld3(F.nrm, p);
        const float *bd = F.band + 5 * p;
        const float3 gA = ld3(F.gaccA, p), gB = ld3(F.gaccB, p);
        s[0][i] = make_float4(a.x, a.y, a.z, bb.x);
        s[1][i] = make_float4(bb.y, bb.z, al.x, al.y);
        s[2][i] = make_float4(al.z, n.x, n.y, n.z);
        s[3][i] = make_float4(bd[0], bd[1], bd[2], bd[3]);
        s[4][i] = make_float4(bd[4], gA.x, gA.y, gA.z);
        s[5][i] = make_float4(F.gwA[p], gB.x, gB.y, gB.z);
        s[6][i] = make_float4(F.gwB[p], 0.f, 0.f, 0.f);
    }
    __syncthreads();
    const int lx = threadIdx.x & (SF_T - 1), ly = threadIdx.x / SF_T;
    const int qx = x0 + lx, qy = y0 + ly;
    if (qx >= W || qy >= H) return;
    const int q0 = (ly + HALF) * E + lx + HALF;
    auto unpack = [&](int k, float3 &cA, float3 &cB, float3 &ca, float3 &cn) {
        const float4 P0 = s[0][k], P1 = s[1][k], P2 = s[2][k];
        cA = make_float3(P0.x, P0.y, P0.z);
        cB = make_float3(P0.w, P1.x, P1.y);
        ca = make_float3(P1.z, P1.w, P2.x);
        cn = make_float3(P2.y, P2.z, P2.w);
    };
    float3 qA, qB, qa, qn;
    unpack(q0, qA, qB, qa, qn);
    const float4 qb03 = s[3][q0], qb4g = s[4][q0], qhg = s[5][q0];
    const float qhB = s[6][q0].x;
    const float3 qgA = make_float3(qb4g.y, qb4g.z, qb4g.w), qgB = make_float3(qhg.y, qhg.z, qhg.w);
    const float qhA = qhg.x;
    float3 dA = make_float3(0.f, 0.f, 0.f), dB = dA, da = dA, dn = dA;
    const float inv_h2 = 1.f / (float)(HALF * HALF);
    for (int dy = -HALF; dy <= HALF; ++dy)
    {
        int pys[3], npy = 0;
        {
            const int xs[3] = {qy, qy > 0 ? -qy : INT_MIN, qy <= H - 2 ? 2 * H - 2 - qy : INT_MIN};
            for (int k = 0; k < 3; ++k)
                if (xs[k] != INT_MIN && xs[k] - dy >= 0 && xs[k] - dy < H) pys[npy++] = xs[k] - dy;
        }
        for (int dx = -HALF; dx <= HALF; ++dx)
        {
            const float pd = (float)(dx * dx + dy * dy) * inv_h2;
            // own term: tap at q + d
            {
                const int k = q0 + dy * E + dx;
                float3 iA, iB, ia, in;
                unpack(k, iA, iB, ia, in);
                const float eA = dist2(qA, iA), eB = dist2(qB, iB), ea = dist2(qa, ia), en = dist2(qn, in);
                const float sh = -qb03.z * ea - qb03.w * en - qb4g.x * pd;
                const float wA = __expf(-qb03.x * eA + sh), wB = __expf(-qb03.y * eB + sh);
                const float uA = wA * (qgA.x * iA.x + qgA.y * iA.y + qgA.z * iA.z + qhA);
                const float uB = wB * (qgB.x * iB.x + qgB.y * iB.y + qgB.z * iB.z + qhB);
                const float tA = -2.f * qb03.x * uA, tB = -2.f * qb03.y * uB;
                const float ta = -2.f * qb03.z * (uA + uB), tn = -2.f * qb03.w * (uA + uB);
                dA.x += tA * (qA.x - iA.x); dA.y += tA * (qA.y - iA.y); dA.z += tA * (qA.z - iA.z);
                dB.x += tB * (qB.x - iB.x); dB.y += tB * (qB.y - iB.y); dB.z += tB * (qB.z - iB.z);
                da.x += ta * (qa.x - ia.x); da.y += ta * (qa.y - ia.y); da.z += ta * (qa.z - ia.z);
                dn.x += tn * (qn.x - in.x); dn.y += tn * (qn.y - in.y); dn.z += tn * (qn.z - in.z);
            }
            // tap terms: centers p with mirror(p + d) = q
            int pxs[3], npx = 0;
            {
                const int xs[3] = {qx, qx > 0 ? -qx : INT_MIN, qx <= W - 2 ? 2 * W - 2 - qx : INT_MIN};
                for (int k = 0; k < 3; ++k)
                    if (xs[k] != INT_MIN && xs[k] - dx >= 0 && xs[k] - dx < W) pxs[npx++] = xs[k] - dx;
            }
            for (int a = 0; a < npy; ++a)
                for (int bq = 0; bq < npx; ++bq)
                {
                    const int k = (pys[a] - y0 + HALF) * E + (pxs[bq] - x0 + HALF);
                    float3 pA, pB, pa, pn;
                    unpack(k, pA, pB, pa, pn);
                    const float4 b03 = s[3][k], b4g = s[4][k], hg = s[5][k];
                    const float phB = s[6][k].x;
                    const float3 gA = make_float3(b4g.y, b4g.z, b4g.w), gB = make_float3(hg.y, hg.z, hg.w);
                    const float eA = dist2(pA, qA), eB = dist2(pB, qB), ea = dist2(pa, qa), en = dist2(pn, qn);
                    const float sh = -b03.z * ea - b03.w * en - b4g.x * pd;
                    const float wA = __expf(-b03.x * eA + sh), wB = __expf(-b03.y * eB + sh);
                    const float uA = wA * (gA.x * qA.x + gA.y * qA.y + gA.z * qA.z + hg.x);
                    const float uB = wB * (gB.x * qB.x + gB.y * qB.y + gB.z * qB.z + phB);
                    const float tA = 2.f * b03.x * uA, tB = 2.f * b03.y * uB;
                    const float ta = 2.f * b03.z * (uA + uB), tn = 2.f * b03.w * (uA + uB);
                    dA.x += wA * gA.x + tA * (pA.x - qA.x);
                    dA.y += wA * gA.y + tA * (pA.y - qA.y);
                    dA.z += wA * gA.z + tA * (pA.z - qA.z);
                    dB.x += wB * gB.x + tB * (pB.x - qB.x);
                    dB.y += wB * gB.y + tB * (pB.y - qB.y);
                    dB.z += wB * gB.z + tB * (pB.z - qB.z);
                    da.x += ta * (pa.x - qa.x); da.y += ta * (pa.y - qa.y); da.z += ta * (pa.z - qa.z);
                    dn.x += tn * (pn.x - qn.x); dn.y += tn * (pn.y - qn.y); dn.z += tn * (pn.z - qn.z);
                }
        }
    }
    const size_t c = (size_t)qy * W + qx;
    if (F.gia) st3(F.gia, c, dA);
    if (F.gib) st3(F.gib, c, dB);
    if (F.galb) st3(F.galb, c, da);
    if (F.gnrm) st3(F.gnrm, c, dn);
}

// ---------------------------------------------------------------------------------------------
static int div_up(int a, int b) { return (a + b - 1) / b; }

static SfBatch sf_batch(int n, const SfFrame *fr, int win)
{
    SfBatch b{};
    int tx[XDN_MAX_FRAMES], ty[XDN_MAX_FRAMES];
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        tx[i] = div_up(fr[i].W, SF_T);
        ty[i] = div_up(fr[i].H, SF_T);
    }
    b.map = make_tilemap(n, tx, ty);
    b.half = win / 2;
    return b;
}

hipError_t launch_sf_forward(int n, const SfFrame *fr, int win, hipStream_t st)
{
    SfBatch b = sf_batch(n, fr, win);
    if (b.map.total == 0) return hipSuccess;
    if (b.half == 5)
        hipLaunchKernelGGL((sf_tile<5, SF_FWD>), dim3(b.map.total), dim3(256), 0, st, b);
    else
        hipLaunchKernelGGL((sf_generic<SF_FWD>), dim3(b.map.total), dim3(256), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_sf_grad_band(int n, const SfFrame *fr, int win, hipStream_t st)
{
    SfBatch b = sf_batch(n, fr, win);
    if (b.map.total == 0) return hipSuccess;
    if (b.half == 5)
        hipLaunchKernelGGL((sf_tile<5, SF_GRAD>), dim3(b.map.total), dim3(256), 0, st, b);
    else
        hipLaunchKernelGGL((sf_generic<SF_GRAD>), dim3(b.map.total), dim3(256), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_sf_grad_inputs(int n, const SfFrame *fr, int win, hipStream_t st)
{
    SfBatch b = sf_batch(n, fr, win);
    if (b.map.total == 0) return hipSuccess;
    switch (b.half)
    {
    case 1: hipLaunchKernelGGL((sf_grad_inputs_tile<1>), dim3(b.map.total), dim3(256), 0, st, b); break;
    case 2: hipLaunchKernelGGL((sf_grad_inputs_tile<2>), dim3(b.map.total), dim3(256), 0, st, b); break;
    case 3: hipLaunchKernelGGL((sf_grad_inputs_tile<3>), dim3(b.map.total), dim3(256), 0, st, b); break;
    case 4: hipLaunchKernelGGL((sf_grad_inputs_tile<4>), dim3(b.map.total), dim3(256), 0, st, b); break;
    case 5: hipLaunchKernelGGL((sf_grad_inputs_tile<5>), dim3(b.map.total), dim3(256), 0, st, b); break;
    case 6: hipLaunchKernelGGL((sf_grad_inputs_tile<6>), dim3(b.map.total), dim3(256), 0, st, b); break;
    case 7: hipLaunchKernelGGL((sf_grad_inputs_tile<7>), dim3(b.map.total), dim3(256), 0, st, b); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int filter_train_tiles(int H, int W) { return div_up(W, SF_T) * div_up(H, SF_T); }

hipError_t launch_filter_train(int n, const TrainFrame *fr, int win, hipStream_t st)
{
    if (win / 2 != 5) return hipErrorInvalidValue;
    TrBatch b{};
    int tx[XDN_MAX_FRAMES], ty[XDN_MAX_FRAMES];
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        tx[i] = div_up(fr[i].W, SF_T);
        ty[i] = div_up(fr[i].H, SF_T);
    }
    b.map = make_tilemap(n, tx, ty);
    if (b.map.total == 0) return hipSuccess;
    hipLaunchKernelGGL((sf_train_tile<5>), dim3(b.map.total), dim3(256), 0, st, b);
    return hipGetLastError();
}

}  // namespace xdn
