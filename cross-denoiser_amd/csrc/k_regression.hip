// Cross-regression on gfx950: input variance, per-center weighted ridge regression and block
// reconstruction (the cross_regression op, regression.cu.cc:244-500).
//
// Data layout in HBM: images are NHWC float32 with 3 channels (12 B/px) exactly as the op
// boundary hands them over. The regression coefficients are an internal workspace laid out as
// [B][H/s][W/s][2 halves][6 coeffs][3 channels] float32 (144 B per center).
//
// Kernels
//   variance_kernel    3x3 clamped neighbourhood variance (EstInputVarianceKernel :244-276),
//                      HBM-bound, 1 px/thread. Optionally fuses log1p of the raw colour
//                      (scripts/main.py:207-212) and writes the log image.
//   regression_rows    fast path (win 17, sparsity 4): one workgroup = 64 centers of one center
//                      row, both halves; the 17 tap rows stream through a double-buffered LDS
//                      row (4 float4 planes per pixel) and each (center, half) is served by two
//                      lanes that split the 17 tap columns 9/8 and are reduced with a lane swap.
//                      61 FP32 accumulators per lane; 6x6 Cholesky per channel spread over the
//                      lane pair. VALU-bound.
//   recon_tile         fast path (win 17, sparsity 4, recon_scale 1): 16x16 pixels per
//                      workgroup; the <= 8x8 centers the tile touches are staged in LDS (both
//                      halves' coefficients and features), one pixel per lane, both halves.
//   regression_generic / recon_generic: any win_size / sparsity / recon_scale, direct loads.
#include "xdn_common.h"
#include "xdn_kernels.h"
#include "xdn_regression.h"

namespace xdn {

// ---------------------------------------------------------------------------------------------
// Variance (and optional log1p prologue)
// ---------------------------------------------------------------------------------------------
struct VarFrame {
    const float *in;   // y (already log) or raw colour when log_in is set
    float *y_out;      // log1p(colour) when log_in is set, else unused
    float *var;        // [H][W]
    int H, W;
};
struct VarBatch {
    TileMap map;
    VarFrame f[XDN_MAX_FRAMES];
    int log_in;
};

constexpr int VT_X = 32, VT_Y = 16;

// log1p(x) = ln(u) * x / (u - 1) with u = fl(1 + x) (Goldberg 1991, "What every computer scientist
// should know about floating-point arithmetic", Theorem 4): the rounding error of 1 + x cancels in
// the ratio, so a few-ulp ln(u) (v_log_f32 is log2 to ~1 ulp) gives a few-ulp log1p at ~10
// instructions instead of the libm polynomial.
__device__ __forceinline__ float log1p_acc(float x)
{
    const float u = 1.f + x;
    if (u == 1.f) return x;
    if (u == __builtin_inff()) return u;
    return (__builtin_amdgcn_logf(u) * 0.693147180559945309f) * (x * fast_rcp(u - 1.f));
}

// 32x16 pixels per workgroup, two per lane; the tile and its 1-px halo (zero outside the image)
// are loaded with every load of a lane in flight at once. The neighbour sum runs over all eight
// neighbours in the reference's order: the zeros of out-of-image taps leave it unchanged, and the
// count of in-image neighbours follows from the position.
__global__ __launch_bounds__(256) void variance_kernel(VarBatch b)
{
    constexpr int EX = VT_X + 2, EY = VT_Y + 2, NS = (EX * EY + 255) / 256;
    __shared__ float s[3][EY][EX];
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const VarFrame &F = b.f[L.f];
    const int H = F.H, W = F.W;
    const int x0 = L.tx * VT_X, y0 = L.ty * VT_Y;
    float3 v[NS];
#pragma unroll
    for (int u = 0; u < NS; ++u)
    {
        const int i = threadIdx.x + 256 * u;
        const int ly = i / EX, lx = i - ly * EX;
        const int gx = x0 + lx - 1, gy = y0 + ly - 1;
        v[u] = make_float3(0.f, 0.f, 0.f);
        if (i < EX * EY && gx >= 0 && gy >= 0 && gx < W && gy < H) v[u] = ld3(F.in, (size_t)gy * W + gx);
    }
#pragma unroll
    for (int u = 0; u < NS; ++u)
    {
        const int i = threadIdx.x + 256 * u;
        if (i >= EX * EY) continue;
        const int ly = i / EX, lx = i - ly * EX;
        const int gx = x0 + lx - 1, gy = y0 + ly - 1;
        float3 t = v[u];
        if (b.log_in && gx >= 0 && gy >= 0 && gx < W && gy < H)
        {
            t = make_float3(log1p_acc(t.x), log1p_acc(t.y), log1p_acc(t.z));
            if (lx >= 1 && ly >= 1 && lx <= VT_X && ly <= VT_Y) st3(F.y_out, (size_t)gy * W + gx, t);
        }
        s[0][ly][lx] = t.x;
        s[1][ly][lx] = t.y;
        s[2][ly][lx] = t.z;
    }
    __syncthreads();
    const int lx = threadIdx.x % VT_X;
#pragma unroll
    for (int r = 0; r < 2; ++r)
    {
        const int ly = threadIdx.x / VT_X + 8 * r;
        const int cx = x0 + lx, cy = y0 + ly;
        if (cx >= W || cy >= H) continue;
        const int n = (1 + (cx > 0) + (cx < W - 1)) * (1 + (cy > 0) + (cy < H - 1)) - 1;
        float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx)
            {
                if ((dx | dy) == 0) continue;
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[k] += s[k][ly + 1 + dy][lx + 1 + dx];
            }
        const float inv = 1.f / (float)n;
        float var = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            const float d = acc[k] * inv - s[k][ly + 1][lx + 1];
            var += d * d;
        }
        F.var[(size_t)cy * W + cx] = var;
    }
}

// ---------------------------------------------------------------------------------------------
// Regression
// ---------------------------------------------------------------------------------------------
struct RegFrame {
    const float *y1, *y2, *var1, *var2, *alb, *nrm;
    float *beta;  // [Hs][Ws][2][18]
    int H, W;
    int nfull;    // fast path with paired row tails: full 64-center blocks per center row, else 0
};
struct RegBatch {
    TileMap map;
    RegFrame f[XDN_MAX_FRAMES];
    int half_win, sparsity;  // generic kernel only
};

__device__ __forceinline__ float swap_add(float v) { return v + __shfl_xor(v, 1); }

__device__ __forceinline__ void reduce_pair(RegAcc &a)
{
    a.s00 = swap_add(a.s00); a.s03 = swap_add(a.s03); a.s04 = swap_add(a.s04); a.s05 = swap_add(a.s05);
    a.s33 = swap_add(a.s33); a.s34 = swap_add(a.s34); a.s35 = swap_add(a.s35); a.s44 = swap_add(a.s44);
    a.s45 = swap_add(a.s45); a.s55 = swap_add(a.s55);
#pragma unroll
    for (int c = 0; c < 3; ++c)
    {
        a.c01[c] = swap_add(a.c01[c]); a.c02[c] = swap_add(a.c02[c]); a.c11[c] = swap_add(a.c11[c]);
        a.c12[c] = swap_add(a.c12[c]); a.c22[c] = swap_add(a.c22[c]); a.c13[c] = swap_add(a.c13[c]);
        a.c14[c] = swap_add(a.c14[c]); a.c15[c] = swap_add(a.c15[c]); a.c23[c] = swap_add(a.c23[c]);
        a.c24[c] = swap_add(a.c24[c]); a.c25[c] = swap_add(a.c25[c]);
        a.x0[c] = swap_add(a.x0[c]); a.x1[c] = swap_add(a.x1[c]); a.x2[c] = swap_add(a.x2[c]);
        a.x3[c] = swap_add(a.x3[c]); a.x4[c] = swap_add(a.x4[c]); a.x5[c] = swap_add(a.x5[c]);
    }
}

constexpr int RG_CX = 64;       // centers per workgroup
constexpr int RG_S = 4;         // sparsity of the fast path
constexpr int RG_H = 8;         // half window of the fast path
constexpr int RG_ROW = RG_CX * RG_S + 2 * RG_H;

// Launch shape: one block per 64 centers of a center row; where a row's last block would hold at
// most 32 centers (1920 px: 480 centers = 7.5 blocks), the tails of two consecutive rows share one
// block instead (paired tail: wave w serves row 2p + (w & 1), half w >> 1, 32 centers each, over the
// 21 tap rows the two center rows span), so a 1080p frame is 2,025 blocks instead of 2,160: two
// rounds of the 1,024 resident at a time (four per CU) instead of three. Tile map per row pair p:
// nfull blocks of row 2p, nfull of row 2p + 1, then the paired tail. Each lane sums its taps in the
// same order either way.
__global__ __launch_bounds__(256) void regression_rows(RegBatch b)
{
    // [buffer][plane][pixel]; planes: (y1, var1), (y2, var2), (albedo, std1), (normal, std2)
    __shared__ float4 lds[2][4][RG_ROW];
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const RegFrame &F = b.f[L.f];
    const int H = F.H, W = F.W, Ws = W / RG_S, Hs = H / RG_S;
    const int t = threadIdx.x;
    int by, bx0;
    bool paired = false;
    if (F.nfull)
    {
        const int c = L.tx, nf = F.nfull;
        paired = c == 2 * nf;
        by = 2 * L.ty + (c >= nf && !paired);
        bx0 = (paired ? nf : (c >= nf ? c - nf : c)) * RG_CX;
        if (by >= Hs) return;  // odd row count: the last pair has no second row (whole block)
    }
    else
    {
        by = L.ty;
        bx0 = L.tx * RG_CX;
    }
    const int half = t >> 7;           // wave-uniform: waves 0-1 half A, waves 2-3 half B
    const int j = t & 127;
    const int o = paired ? (j >> 6) : 0;  // center row offset of this wave (paired tail only)
    const int ci = paired ? ((j & 63) >> 1) : (j >> 1), split = j & 1;
    const int row_c = by + o;          // this lane's center row
    const int bx = bx0 + ci;
    const bool active = bx < Ws && row_c < Hs;
    const int cx = (active ? bx : Ws - 1) * RG_S, cy = min(row_c, Hs - 1) * RG_S;
    const int px0 = bx0 * RG_S - RG_H;  // pixel x of LDS column 0

    const float *feat = half ? F.y2 : F.y1;
    const float *fvar = half ? F.var2 : F.var1;
    const size_t c = (size_t)cy * W + cx;
    const float3 cw = ld3(feat, c), ca = ld3(F.alb, c), cn = ld3(F.nrm, c);
    const float cvar = fvar[c];
    const float cvar_e = cvar + 1e-4f;
    const float cstd_e = sqrtf(cvar) + 1e-2f;

    RegAcc acc;
    acc.zero();

    const int ncols = min(RG_ROW, W + 2 * RG_H - bx0 * RG_S);  // columns actually needed
    // a wave whose centers all lie past the row end (or on a missing second row) only stages rows
    // and meets the barriers; its lanes' results are never stored
    const bool wave_busy = __builtin_amdgcn_readfirstlane(
        (bx0 + (paired ? 0 : (j & ~63) / 2)) < Ws && row_c < Hs);
    const int own = half, oth = 1 - half;
    // first tap column of this lane: split 0 covers dx in [-8, 0], split 1 covers [1, 8]
    const int col0 = ci * RG_S + (split ? RG_H + 1 : 0);
    const int ntap = split ? RG_H : RG_H + 1;
    const int nrows = 2 * RG_H + 1 + (paired ? RG_S : 0);

    for (int r = 0; r < nrows; ++r)
    {
        const int buf = r & 1;
        const int gy = mirror(by * RG_S + r - RG_H, H);
        const size_t row = (size_t)gy * W;
        for (int i = t; i < ncols; i += 256)
        {
            const size_t p = row + mirror(px0 + i, W);
            const float3 a = ld3(F.y1, p), bb = ld3(F.y2, p), al = ld3(F.alb, p), nr = ld3(F.nrm, p);
            const float v1 = F.var1[p], v2 = F.var2[p];
            lds[buf][0][i] = make_float4(a.x, a.y, a.z, v1);
            lds[buf][1][i] = make_float4(bb.x, bb.y, bb.z, v2);
            lds[buf][2][i] = make_float4(al.x, al.y, al.z, sqrtf(v1));
            lds[buf][3][i] = make_float4(nr.x, nr.y, nr.z, sqrtf(v2));
        }
        __syncthreads();
        if (!wave_busy || r < o * RG_S || r > o * RG_S + 2 * RG_H) continue;
#pragma unroll
        for (int k = 0; k < RG_H + 1; ++k)
        {
            if (k < ntap)
            {
                const int col = col0 + k;
                const float4 P = lds[buf][own][col];
                const float4 T = lds[buf][oth][col];
                const float4 A = lds[buf][2][col];
                const float4 N = lds[buf][3][col];
                const float istd = half ? N.w : A.w;
                const float3 dw = make_float3(P.x - cw.x, P.y - cw.y, P.z - cw.z);
                const float w = reg_weight(dw, cvar_e, P.w);
                const float inv = fast_rcp(cstd_e + istd);
                acc.add(w, inv, dw, make_float3(A.x - ca.x, A.y - ca.y, A.z - ca.z),
                        make_float3(N.x - cn.x, N.y - cn.y, N.z - cn.z), make_float3(T.x, T.y, T.z));
            }
        }
    }
    reduce_pair(acc);
    if (!active) return;
    // Lane pair solves the three channels: split 0 -> channels 0 and 1, split 1 -> channel 2.
    float *out = F.beta + (((size_t)row_c * Ws + bx) * 2 + half) * 18;
    {
        Sys6 s0, s2;
        acc.system(0, s0);
        acc.system(2, s2);
        Sys6 s;
#pragma unroll
        for (int k = 0; k < 21; ++k) s.a[k] = split ? s2.a[k] : s0.a[k];
#pragma unroll
        for (int k = 0; k < 6; ++k) s.x[k] = split ? s2.x[k] : s0.x[k];
        float be[6];
        solve6(s, be);
        const int ch = split ? 2 : 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) out[k * 3 + ch] = be[k];
    }
    if (split == 0)
    {
        Sys6 s1;
        acc.system(1, s1);
        float be[6];
        solve6(s1, be);
#pragma unroll
        for (int k = 0; k < 6; ++k) out[k * 3 + 1] = be[k];
    }
}

// Any window / sparsity: one lane per (center, half), direct loads.
__global__ __launch_bounds__(256) void regression_generic(RegBatch b)
{
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const RegFrame &F = b.f[L.f];
    const int S = b.sparsity, HW = b.half_win;
    const int H = F.H, W = F.W, Ws = W / S, Hs = H / S;
    const int half = threadIdx.x >> 7;
    const int bx = L.tx * 16 + (threadIdx.x & 15);
    const int by = L.ty * 8 + ((threadIdx.x & 127) >> 4);
    if (bx >= Ws || by >= Hs) return;
    const float *feat = half ? F.y2 : F.y1;
    const float *tgt = half ? F.y1 : F.y2;
    const float *fvar = half ? F.var2 : F.var1;
    const int cx = bx * S, cy = by * S;
    const size_t c = (size_t)cy * W + cx;
    const float3 cw = ld3(feat, c), ca = ld3(F.alb, c), cn = ld3(F.nrm, c);
    const float cvar = fvar[c], cvar_e = cvar + 1e-4f, cstd_e = sqrtf(cvar) + 1e-2f;
    RegAcc acc;
    acc.zero();
    for (int iy = cy - HW; iy <= cy + HW; ++iy)
        for (int ix = cx - HW; ix <= cx + HW; ++ix)
        {
            const size_t i = (size_t)mirror(iy, H) * W + mirror(ix, W);
            const float3 iw = ld3(feat, i), ia = ld3(F.alb, i), in = ld3(F.nrm, i), it = ld3(tgt, i);
            const float ivar = fvar[i];
            const float3 dw = make_float3(iw.x - cw.x, iw.y - cw.y, iw.z - cw.z);
            const float w = reg_weight(dw, cvar_e, ivar);
            const float inv = fast_rcp(cstd_e + sqrtf(ivar));
            acc.add(w, inv, dw, make_float3(ia.x - ca.x, ia.y - ca.y, ia.z - ca.z),
                    make_float3(in.x - cn.x, in.y - cn.y, in.z - cn.z), it);
        }
    float *out = F.beta + (((size_t)by * Ws + bx) * 2 + half) * 18;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
    {
        Sys6 s;
        acc.system(ch, s);
        float be[6];
        solve6(s, be);
#pragma unroll
        for (int k = 0; k < 6; ++k) out[k * 3 + ch] = be[k];
    }
}

// ---------------------------------------------------------------------------------------------
// Block reconstruction
// ---------------------------------------------------------------------------------------------
struct RecFrame {
    const float *beta;  // [Hs][Ws][2][18]
    const float *y1, *y2, *var1, *var2, *alb, *nrm;
    float *outa, *outb;  // [H/rs][W/rs][3]
    int H, W;
};
struct RecBatch {
    TileMap map;
    RecFrame f[XDN_MAX_FRAMES];
    int half_win, sparsity, recon_scale;  // generic kernel only
};

constexpr int RC_T = 16;                 // pixels per tile side
constexpr int RC_C = 8;                  // centers per tile side (fast path)
constexpr int RC_STRIDE = 13;            // float4 per staged center

// Per-center contribution for one half: pred = beta . delta with delta = pixel - center.
__device__ __forceinline__ void recon_term(const float4 *cb, float3 dwv, float inv, float3 d2, float3 n,
                                           float w, float3 &acc)
{
    // cb: 18 floats of beta [k][ch] as 4.5 float4 starting at a 16-B boundary or at +8 B; we
    // read through a float pointer to keep it simple.
    const float *bt = reinterpret_cast<const float *>(cb);
    const float d1x = dwv.x * inv, d1y = dwv.y * inv, d1z = dwv.z * inv;
    const float px = bt[0] + bt[3] * d1x + bt[6] * d2.x + bt[9] * n.x + bt[12] * n.y + bt[15] * n.z;
    const float py = bt[1] + bt[4] * d1y + bt[7] * d2.y + bt[10] * n.x + bt[13] * n.y + bt[16] * n.z;
    const float pz = bt[2] + bt[5] * d1z + bt[8] * d2.z + bt[11] * n.x + bt[14] * n.y + bt[17] * n.z;
    acc.x += w * px;
    acc.y += w * py;
    acc.z += w * pz;
}

// The same prediction with the center's terms folded into the staged record: with
// K0 = beta0 - beta2 . a_c - beta3 . n_c (per channel), beta . delta = K0 + beta1 d1 + beta2 a_p +
// beta3 . n_p, so a (pixel, center) pair needs no albedo or normal difference. bt: K0 in place of
// beta0, then beta1..beta5 as in recon_term; ca, cn: the pixel's albedo and normal.
__device__ __forceinline__ void recon_term_k0(const float *bt, float3 dwv, float inv, float3 ca, float3 cn, float w,
                                              float3 &acc)
{
    const float d1x = dwv.x * inv, d1y = dwv.y * inv, d1z = dwv.z * inv;
    const float px = bt[0] + bt[3] * d1x + bt[6] * ca.x + bt[9] * cn.x + bt[12] * cn.y + bt[15] * cn.z;
    const float py = bt[1] + bt[4] * d1y + bt[7] * ca.y + bt[10] * cn.x + bt[13] * cn.y + bt[16] * cn.z;
    const float pz = bt[2] + bt[5] * d1z + bt[8] * ca.z + bt[11] * cn.x + bt[14] * cn.y + bt[17] * cn.z;
    acc.x += w * px;
    acc.y += w * py;
    acc.z += w * pz;
}

__global__ __launch_bounds__(256) void recon_tile(RecBatch b)
{
    // staged center record (13 float4): beta A (18) + beta B (18) = 9 float4,
    // (y1, var1), (y2, var2), (alb, std1), (nrm, std2)
    __shared__ float4 cs[RC_C * RC_C][RC_STRIDE];
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const RecFrame &F = b.f[L.f];
    const int H = F.H, W = F.W, Ws = W / RG_S;
    const int x0 = L.tx * RC_T, y0 = L.ty * RC_T;
    // tile center window [clo, clo+RC_C) x [rlo, rlo+RC_C), a superset of every pixel's range
    const int clo = max(0, (x0 - RG_H + RG_S - 1 + 4 * RG_S) / RG_S - 4);
    const int rlo = max(0, (y0 - RG_H + RG_S - 1 + 4 * RG_S) / RG_S - 4);
    const int chi = min((W - 1) / RG_S, (x0 + RC_T - 1 + RG_H) / RG_S);
    const int rhi = min((H - 1) / RG_S, (y0 + RC_T - 1 + RG_H) / RG_S);
    const int ncx = chi - clo + 1, ncy = rhi - rlo + 1;
    for (int i = threadIdx.x; i < ncx * ncy * 9; i += 256)
    {
        const int cidx = i / 9, q = i - cidx * 9;
        const int ry = cidx / ncx, rx = cidx - ry * ncx;
        const float4 *src = reinterpret_cast<const float4 *>(F.beta + ((size_t)(rlo + ry) * Ws + (clo + rx)) * 36);
        cs[ry * RC_C + rx][q] = src[q];
    }
    for (int i = threadIdx.x; i < ncx * ncy; i += 256)
    {
        const int ry = i / ncx, rx = i - ry * ncx;
        const size_t p = (size_t)((rlo + ry) * RG_S) * W + (size_t)((clo + rx) * RG_S);
        const float3 a = ld3(F.y1, p), bb = ld3(F.y2, p), al = ld3(F.alb, p), nr = ld3(F.nrm, p);
        const float v1 = F.var1[p], v2 = F.var2[p];
        float4 *d = cs[ry * RC_C + rx];
        d[9] = make_float4(a.x, a.y, a.z, v1);
        d[10] = make_float4(bb.x, bb.y, bb.z, v2);
        d[11] = make_float4(al.x, al.y, al.z, sqrtf(v1));
        d[12] = make_float4(nr.x, nr.y, nr.z, sqrtf(v2));
    }
    __syncthreads();
    // fold each center's albedo and normal into its constant term, per half and channel
    for (int i = threadIdx.x; i < ncx * ncy * 2; i += 256)
    {
        const int cidx = i >> 1, hb = i & 1;
        const int ry = cidx / ncx, rx = cidx - ry * ncx;
        float4 *d = cs[ry * RC_C + rx];
        float *bt = reinterpret_cast<float *>(d) + 18 * hb;
        const float4 A = d[11], N = d[12];
        const float av[3] = {A.x, A.y, A.z};
#pragma unroll
        for (int ch = 0; ch < 3; ++ch)
            bt[ch] = bt[ch] - bt[6 + ch] * av[ch] - (bt[9 + ch] * N.x + bt[12 + ch] * N.y + bt[15 + ch] * N.z);
    }
    __syncthreads();
    const int cx = x0 + (threadIdx.x & 15), cy = y0 + (threadIdx.x >> 4);
    if (cx >= W || cy >= H) return;
    const size_t c = (size_t)cy * W + cx;
    const float3 c1 = ld3(F.y1, c), c2 = ld3(F.y2, c), ca = ld3(F.alb, c), cn = ld3(F.nrm, c);
    const float v1 = F.var1[c], v2 = F.var2[c];
    const float v1e = v1 + 1e-4f, v2e = v2 + 1e-4f;
    const float s1e = sqrtf(v1) + 1e-2f, s2e = sqrtf(v2) + 1e-2f;
    const int sx = max(0, idivup(cx - RG_H, RG_S)), ex = min((W - 1) / RG_S, (cx + RG_H) / RG_S);
    const int sy = max(0, idivup(cy - RG_H, RG_S)), ey = min((H - 1) / RG_S, (cy + RG_H) / RG_S);
    float3 accA = make_float3(0.f, 0.f, 0.f), accB = accA;
    float sumA = 0.f, sumB = 0.f;
    for (int iy = sy; iy <= ey; ++iy)
        for (int ix = sx; ix <= ex; ++ix)
        {
            const float4 *rec = cs[(iy - rlo) * RC_C + (ix - clo)];
            const float4 P1 = rec[9], P2 = rec[10];
            const float std1 = rec[11].w, std2 = rec[12].w;
            const float3 dwA = make_float3(c1.x - P1.x, c1.y - P1.y, c1.z - P1.z);
            const float3 dwB = make_float3(c2.x - P2.x, c2.y - P2.y, c2.z - P2.z);
            const float wA = reg_weight(dwA, v1e, P1.w), wB = reg_weight(dwB, v2e, P2.w);
            const float *bt = reinterpret_cast<const float *>(rec);
            recon_term_k0(bt, dwA, fast_rcp(s1e + std1), ca, cn, wA, accA);
            recon_term_k0(bt + 18, dwB, fast_rcp(s2e + std2), ca, cn, wB, accB);
            sumA += wA;
            sumB += wB;
        }
    const float iA = 1.f / fmaxf(sumA, 1e-7f), iB = 1.f / fmaxf(sumB, 1e-7f);
    st3(F.outa, c, make_float3(fmaxf(0.f, accA.x * iA), fmaxf(0.f, accA.y * iA), fmaxf(0.f, accA.z * iA)));
    st3(F.outb, c, make_float3(fmaxf(0.f, accB.x * iB), fmaxf(0.f, accB.y * iB), fmaxf(0.f, accB.z * iB)));
}

__global__ __launch_bounds__(256) void recon_generic(RecBatch b)
{
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const RecFrame &F = b.f[L.f];
    const int S = b.sparsity, HW = b.half_win, RS = b.recon_scale;
    const int H = F.H, W = F.W, Ws = W / S, Hr = H / RS, Wr = W / RS;
    const int rx = L.tx * 16 + (threadIdx.x & 15), ry = L.ty * 16 + (threadIdx.x >> 4);
    if (rx >= Wr || ry >= Hr) return;
    const int cx = rx * RS, cy = ry * RS;
    const size_t c = (size_t)cy * W + cx;
    const float3 c1 = ld3(F.y1, c), c2 = ld3(F.y2, c), ca = ld3(F.alb, c), cn = ld3(F.nrm, c);
    const float v1 = F.var1[c], v2 = F.var2[c];
    const float v1e = v1 + 1e-4f, v2e = v2 + 1e-4f;
    const float s1e = sqrtf(v1) + 1e-2f, s2e = sqrtf(v2) + 1e-2f;
    const int sx = max(0, idivup(cx - HW, S)), ex = min((W - 1) / S, (cx + HW) / S);
    const int sy = max(0, idivup(cy - HW, S)), ey = min((H - 1) / S, (cy + HW) / S);
    float3 accA = make_float3(0.f, 0.f, 0.f), accB = accA;
    float sumA = 0.f, sumB = 0.f;
    for (int iy = sy; iy <= ey; ++iy)
        for (int ix = sx; ix <= ex; ++ix)
        {
            const float *bt = F.beta + ((size_t)iy * Ws + ix) * 36;
            const size_t i = (size_t)(iy * S) * W + (size_t)(ix * S);
            const float3 p1 = ld3(F.y1, i), p2 = ld3(F.y2, i), ia = ld3(F.alb, i), in = ld3(F.nrm, i);
            const float iv1 = F.var1[i], iv2 = F.var2[i];
            const float3 d2 = make_float3(ca.x - ia.x, ca.y - ia.y, ca.z - ia.z);
            const float3 n = make_float3(cn.x - in.x, cn.y - in.y, cn.z - in.z);
            const float3 dwA = make_float3(c1.x - p1.x, c1.y - p1.y, c1.z - p1.z);
            const float3 dwB = make_float3(c2.x - p2.x, c2.y - p2.y, c2.z - p2.z);
            const float wA = reg_weight(dwA, v1e, iv1), wB = reg_weight(dwB, v2e, iv2);
            recon_term(reinterpret_cast<const float4 *>(bt), dwA, fast_rcp(s1e + sqrtf(iv1)), d2, n, wA, accA);
            recon_term(reinterpret_cast<const float4 *>(bt + 18), dwB, fast_rcp(s2e + sqrtf(iv2)), d2, n, wB, accB);
            sumA += wA;
            sumB += wB;
        }
    const size_t o = (size_t)ry * Wr + rx;
    const float iA = 1.f / fmaxf(sumA, 1e-7f), iB = 1.f / fmaxf(sumB, 1e-7f);
    st3(F.outa, o, make_float3(fmaxf(0.f, accA.x * iA), fmaxf(0.f, accA.y * iA), fmaxf(0.f, accA.z * iA)));
    st3(F.outb, o, make_float3(fmaxf(0.f, accB.x * iB), fmaxf(0.f, accB.y * iB), fmaxf(0.f, accB.z * iB)));
}

// ---------------------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------------------
static int div_up(int a, int b) { return (a + b - 1) / b; }

hipError_t launch_variance(int n, const CrFrame *fr, int which, bool log_in, hipStream_t st)
{
    // which: bit0 -> half A (y1/var1), bit1 -> half B (y2/var2); both halves share one launch
    VarBatch b{};
    int tx[XDN_MAX_FRAMES], ty[XDN_MAX_FRAMES];
    int m = 0;
    for (int i = 0; i < n; ++i)
        for (int h = 0; h < 2; ++h)
        {
            if (!(which & (1 << h))) continue;
            VarFrame &v = b.f[m];
            v.in = log_in ? (h ? fr[i].color2 : fr[i].color) : (h ? fr[i].y2 : fr[i].y1);
            v.y_out = h ? fr[i].y2 : fr[i].y1;
            v.var = h ? fr[i].var2 : fr[i].var1;
            v.H = fr[i].H;
            v.W = fr[i].W;
            tx[m] = div_up(v.W, VT_X);
            ty[m] = div_up(v.H, VT_Y);
            ++m;
        }
    b.map = make_tilemap(m, tx, ty);
    b.log_in = log_in ? 1 : 0;
    if (b.map.total == 0) return hipSuccess;
    hipLaunchKernelGGL(variance_kernel, dim3(b.map.total), dim3(256), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_regression(int n, const CrFrame *fr, int win, int s, hipStream_t st)
{
    RegBatch b{};
    int tx[XDN_MAX_FRAMES], ty[XDN_MAX_FRAMES];
    const bool fast = (win / 2 == RG_H && s == RG_S);
    for (int i = 0; i < n; ++i)
    {
        RegFrame &r = b.f[i];
        r.y1 = fr[i].y1; r.y2 = fr[i].y2; r.var1 = fr[i].var1; r.var2 = fr[i].var2;
        r.alb = fr[i].alb; r.nrm = fr[i].nrm; r.beta = fr[i].beta; r.H = fr[i].H; r.W = fr[i].W;
        const int Ws = r.W / s, Hs = r.H / s;
        // pair the rows' tail blocks where a tail holds at most half a block of centers
        const int rem = Ws % RG_CX;
        r.nfull = (fast && Ws >= RG_CX && rem > 0 && rem <= RG_CX / 2 && Hs >= 2) ? Ws / RG_CX : 0;
        if (r.nfull)
        {
            tx[i] = 2 * r.nfull + 1;
            ty[i] = div_up(Hs, 2);
        }
        else
        {
            tx[i] = fast ? div_up(Ws, RG_CX) : div_up(Ws, 16);
            ty[i] = fast ? Hs : div_up(Hs, 8);
        }
    }
    b.map = make_tilemap(n, tx, ty);
    b.half_win = win / 2;
    b.sparsity = s;
    if (b.map.total == 0) return hipSuccess;
    if (fast)
        hipLaunchKernelGGL(regression_rows, dim3(b.map.total), dim3(256), 0, st, b);
    else
        hipLaunchKernelGGL(regression_generic, dim3(b.map.total), dim3(256), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_recon(int n, const CrFrame *fr, int win, int s, int rs, hipStream_t st)
{
    RecBatch b{};
    int tx[XDN_MAX_FRAMES], ty[XDN_MAX_FRAMES];
    const bool fast = (win / 2 == RG_H && s == RG_S && rs == 1);
    for (int i = 0; i < n; ++i)
    {
        RecFrame &r = b.f[i];
        r.beta = fr[i].beta; r.y1 = fr[i].y1; r.y2 = fr[i].y2; r.var1 = fr[i].var1; r.var2 = fr[i].var2;
        r.alb = fr[i].alb; r.nrm = fr[i].nrm; r.outa = fr[i].outa; r.outb = fr[i].outb;
        r.H = fr[i].H; r.W = fr[i].W;
        tx[i] = div_up(r.W / rs, 16);
        ty[i] = div_up(r.H / rs, 16);
    }
    b.map = make_tilemap(n, tx, ty);
    b.half_win = win / 2;
    b.sparsity = s;
    b.recon_scale = rs;
    if (b.map.total == 0) return hipSuccess;
    if (fast)
        hipLaunchKernelGGL(recon_tile, dim3(b.map.total), dim3(256), 0, st, b);
    else
        hipLaunchKernelGGL(recon_generic, dim3(b.map.total), dim3(256), 0, st, b);
    return hipGetLastError();
}

}  // namespace xdn
