// Temporal screening on gfx950 (TemporalScreening, temporal.cu.cc:57-198) with the 3x3 erosion of
// the reprojection success mask (scripts/main.py:270-276) optionally fused in.
//
// A pixel keeps its reprojected history for a set X in {avg, 1, 2} iff for every channel
// min_prev < max_curr and min_curr < max_prev (strict), the current extrema taken over the
// win_curr x win_curr window with out-of-image taps skipped and the previous ones over win_prev
// (only where the eroded success is >= 1). Outputs are selections of the inputs, so the result is
// exact given identical inputs; the set average is computed as (c1 + c2) * 0.5f in FP32.
//
// screening_tile (win_curr 7, win_prev 1): 32x8 pixels per workgroup; the tile plus a 3-px halo of
// (c1, c2, avg) goes to LDS, then a separable min/max (7-tap column pass in place in LDS, 7-tap row
// pass per pixel) gives the 18 extrema per pixel with 14 instead of 49 taps.
// screening_generic: any window sizes, direct loads.
#include "xdn_common.h"
#include "xdn_kernels.h"

#pragma clang fp contract(off)

namespace xdn {

struct TsBatch {
    TileMap map;
    TsFrame f[XDN_MAX_FRAMES];
    int win_curr, win_prev, fuse;
};
static_assert(sizeof(TsBatch) <= 4096, "kernel argument too large");

constexpr float SCR_BIG = 10000000.f;  // FLT_MAX as redefined by the reference (temporal.cu.cc:29)
constexpr int TS_TX = 32, TS_TY = 8, TS_R = 3;

// Eroded success at (x, y): min over the in-bounds 3x3 neighbourhood of the u8 mask.
__device__ __forceinline__ float eroded(const uint8_t *m, int x, int y, int W, int H)
{
    float r = 1.f;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
        {
            const int gx = x + dx, gy = y + dy;
            if (gx < 0 || gy < 0 || gx >= W || gy >= H) continue;
            r = fminf(r, m[(size_t)gy * W + gx] ? 1.f : 0.f);
        }
    return r;
}

struct MinMax9 {
    float mn[9], mx[9];  // sets avg, 1, 2 x channels
};

// Every global load of the block is issued up front (the pixel's history and the success input with
// the tile staging), so a block makes one round trip to memory instead of three. Staging clamps the
// coordinates into the image: a clamped tap repeats a pixel that is already inside the window, so
// the extrema over the fixed 7-tap windows equal those over the in-image taps the reference visits.
//
// Column pass first: a thread owns whole staged columns (set k, column x: 14 rows) and writes their
// seven-row window minima (rows 0..7) and maxima (rows 8..15) back over them, the 8 windows of a
// column from a prefix/suffix split at row 7 (18 min + 18 max instead of 48 + 48). The row pass then
// reads seven neighbours of each. 22 KB of LDS: seven resident blocks per CU.
__global__ __launch_bounds__(256) void screening_tile(TsBatch b)
{
    constexpr int EX = TS_TX + 2 * TS_R, EY = TS_TY + 2 * TS_R;
    constexpr int EXP = EX;                 // row pitch
    constexpr int NCOL = 9 * EX;            // column jobs
    static_assert(EY == 14 && 2 * TS_TY == 16, "the column pass assumes 7-row windows over 14 rows");
    __shared__ float v[9][2 * TS_TY][EXP];  // avg.xyz, c1.xyz, c2.xyz; then column minima / maxima
    __shared__ uint8_t ms[EY][EX];          // success mask (fused erosion), 1 outside the image
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const TsFrame &F = b.f[L.f];
    const int H = F.H, W = F.W;
    const int x0 = L.tx * TS_TX, y0 = L.ty * TS_TY;
    const int olx = threadIdx.x % TS_TX, oly = threadIdx.x / TS_TX;
    const int cx = x0 + olx, cy = y0 + oly;
    const bool live = cx < W && cy < H;
    const size_t c = live ? (size_t)cy * W + cx : 0;
    // this pixel's history and (unfused) success, in flight during the staging and the passes
    const float3 p = ld3(F.prev, c), p1 = ld3(F.prev1, c), p2 = ld3(F.prev2, c);
    const float rs = b.fuse ? 0.f : F.rsucc[c];
    for (int i = threadIdx.x; i < EX * EY; i += 256)
    {
        const int ly = i / EX, lx = i - ly * EX;
        const int rx = x0 + lx - TS_R, ry = y0 + ly - TS_R;
        const int gx = min(max(rx, 0), W - 1), gy = min(max(ry, 0), H - 1);
        const size_t q = (size_t)gy * W + gx;
        const float3 a = ld3(F.c1, q), bb = ld3(F.c2, q);
        if (b.fuse) ms[ly][lx] = (rx == gx && ry == gy) ? (F.succ_mask[q] ? 1 : 0) : 1;
        v[0][ly][lx] = (a.x + bb.x) * 0.5f; v[1][ly][lx] = (a.y + bb.y) * 0.5f; v[2][ly][lx] = (a.z + bb.z) * 0.5f;
        v[3][ly][lx] = a.x; v[4][ly][lx] = a.y; v[5][ly][lx] = a.z;
        v[6][ly][lx] = bb.x; v[7][ly][lx] = bb.y; v[8][ly][lx] = bb.z;
    }
    __syncthreads();
    // this pixel's current values, read before the column pass overwrites the staged tile
    const float3 a = make_float3(v[3][oly + TS_R][olx + TS_R], v[4][oly + TS_R][olx + TS_R], v[5][oly + TS_R][olx + TS_R]);
    const float3 bb = make_float3(v[6][oly + TS_R][olx + TS_R], v[7][oly + TS_R][olx + TS_R], v[8][oly + TS_R][olx + TS_R]);
    __syncthreads();
    for (int j = threadIdx.x; j < NCOL; j += 256)
    {
        const int k = j / EX, x = j - k * EX;
        float t[EY];
#pragma unroll
        for (int r = 0; r < EY; ++r) t[r] = v[k][r][x];
        // window r covers rows r..r+6: suffix extrema of rows r..6 and prefix extrema of rows 7..r+6
        float sn[7], sx[7], pn[7], px[7];
        sn[6] = sx[6] = t[6];
#pragma unroll
        for (int r = 5; r >= 0; --r) sn[r] = fminf(t[r], sn[r + 1]), sx[r] = fmaxf(t[r], sx[r + 1]);
        pn[0] = px[0] = t[7];
#pragma unroll
        for (int r = 1; r < 7; ++r) pn[r] = fminf(pn[r - 1], t[7 + r]), px[r] = fmaxf(px[r - 1], t[7 + r]);
        v[k][0][x] = sn[0];
        v[k][TS_TY][x] = sx[0];
#pragma unroll
        for (int r = 1; r < 7; ++r)
        {
            v[k][r][x] = fminf(sn[r], pn[r - 1]);
            v[k][TS_TY + r][x] = fmaxf(sx[r], px[r - 1]);
        }
        v[k][7][x] = pn[6];
        v[k][TS_TY + 7][x] = px[6];
    }
    __syncthreads();
    if (!live) return;
    float sc = rs;
    if (b.fuse)
    {
        // 3x3 erosion: min over the in-bounds neighbours (outside the image stored as 1)
        int m = 1;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) m = min(m, (int)ms[oly + TS_R + dy][olx + TS_R + dx]);
        sc = m ? 1.f : 0.f;
        if (F.succ_e) F.succ_e[c] = sc;
    }
    const float3 m = make_float3((a.x + bb.x) * 0.5f, (a.y + bb.y) * 0.5f, (a.z + bb.z) * 0.5f);
    if (sc + 1e-4f < 1.f)
    {
        st3(F.o, c, m);
        st3(F.o1, c, a);
        st3(F.o2, c, bb);
        F.osucc[c] = 0.f;
        return;
    }
    // row pass, folded into the keep tests set by set. win_prev == 1: the previous extrema are
    // the center values when sc >= 1, else the initial (+BIG, -BIG) of the reference
    const bool center_ok = sc >= 1.f;
    const float pv[9] = {p.x, p.y, p.z, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z};
    bool keep[3];
#pragma unroll
    for (int st = 0; st < 3; ++st)
    {
        bool kp = true;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch)
        {
            const int k = 3 * st + ch;
            float rn[2 * TS_R + 1], rx[2 * TS_R + 1];
#pragma unroll
            for (int xx = 0; xx <= 2 * TS_R; ++xx) rn[xx] = v[k][oly][olx + xx], rx[xx] = v[k][TS_TY + oly][olx + xx];
            const float mn = fminf(fminf(fminf(rn[0], rn[1]), fminf(rn[2], rn[3])), fminf(fminf(rn[4], rn[5]), rn[6]));
            const float mx = fmaxf(fmaxf(fmaxf(rx[0], rx[1]), fmaxf(rx[2], rx[3])), fmaxf(fmaxf(rx[4], rx[5]), rx[6]));
            const float pmn = center_ok ? pv[k] : SCR_BIG;
            const float pmx = center_ok ? pv[k] : -SCR_BIG;
            kp = kp && (pmn < mx) && (mn < pmx);
        }
        keep[st] = kp;
    }
    st3(F.o, c, keep[0] ? p : m);
    st3(F.o1, c, keep[1] ? p1 : a);
    st3(F.o2, c, keep[2] ? p2 : bb);
    F.osucc[c] = keep[0] ? 1.f : 0.f;
}

__global__ __launch_bounds__(256) void screening_generic(TsBatch b)
{
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const TsFrame &F = b.f[L.f];
    const int H = F.H, W = F.W;
    const int cx = L.tx * TS_TX + (threadIdx.x % TS_TX), cy = L.ty * TS_TY + threadIdx.x / TS_TX;
    if (cx >= W || cy >= H) return;
    const size_t c = (size_t)cy * W + cx;
    const float sc = b.fuse ? eroded(F.succ_mask, cx, cy, W, H) : F.rsucc[c];
    if (b.fuse && F.succ_e) F.succ_e[c] = sc;
    const int hc = b.win_curr / 2, hp = b.win_prev / 2;
    MinMax9 cur, prv;
#pragma unroll
    for (int k = 0; k < 9; ++k) cur.mn[k] = prv.mn[k] = SCR_BIG, cur.mx[k] = prv.mx[k] = -SCR_BIG;
    if (!(sc + 1e-4f < 1.f))
    {
        for (int iy = cy - hc; iy <= cy + hc; ++iy)
            for (int ix = cx - hc; ix <= cx + hc; ++ix)
            {
                if (ix < 0 || ix >= W || iy < 0 || iy >= H) continue;
                const size_t i = (size_t)iy * W + ix;
                const float3 a = ld3(F.c1, i), bb = ld3(F.c2, i);
                const float t[9] = {(a.x + bb.x) * 0.5f, (a.y + bb.y) * 0.5f, (a.z + bb.z) * 0.5f, a.x, a.y, a.z,
                                    bb.x, bb.y, bb.z};
#pragma unroll
                for (int k = 0; k < 9; ++k) cur.mn[k] = fminf(cur.mn[k], t[k]), cur.mx[k] = fmaxf(cur.mx[k], t[k]);
            }
        for (int iy = cy - hp; iy <= cy + hp; ++iy)
            for (int ix = cx - hp; ix <= cx + hp; ++ix)
            {
                if (ix < 0 || ix >= W || iy < 0 || iy >= H) continue;
                const size_t i = (size_t)iy * W + ix;
                const float si = b.fuse ? eroded(F.succ_mask, ix, iy, W, H) : F.rsucc[i];
                if (si < 1.f) continue;
                const float3 p = ld3(F.prev, i), p1 = ld3(F.prev1, i), p2 = ld3(F.prev2, i);
                const float t[9] = {p.x, p.y, p.z, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z};
#pragma unroll
                for (int k = 0; k < 9; ++k) prv.mn[k] = fminf(prv.mn[k], t[k]), prv.mx[k] = fmaxf(prv.mx[k], t[k]);
            }
    }
    const float3 a = ld3(F.c1, c), bb = ld3(F.c2, c);
    const float3 m = make_float3((a.x + bb.x) * 0.5f, (a.y + bb.y) * 0.5f, (a.z + bb.z) * 0.5f);
    if (sc + 1e-4f < 1.f)
    {
        st3(F.o, c, m);
        st3(F.o1, c, a);
        st3(F.o2, c, bb);
        F.osucc[c] = 0.f;
        return;
    }
    bool keep[3];
#pragma unroll
    for (int s = 0; s < 3; ++s)
    {
        bool k = true;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch)
            k = k && (prv.mn[3 * s + ch] < cur.mx[3 * s + ch]) && (cur.mn[3 * s + ch] < prv.mx[3 * s + ch]);
        keep[s] = k;
    }
    st3(F.o, c, keep[0] ? ld3(F.prev, c) : m);
    st3(F.o1, c, keep[1] ? ld3(F.prev1, c) : a);
    st3(F.o2, c, keep[2] ? ld3(F.prev2, c) : bb);
    F.osucc[c] = keep[0] ? 1.f : 0.f;
}

static int div_up(int a, int b) { return (a + b - 1) / b; }

hipError_t launch_screening(int n, const TsFrame *fr, int win_curr, int win_prev, bool fuse_erosion, hipStream_t st)
{
    TsBatch b{};
    int tx[XDN_MAX_FRAMES], ty[XDN_MAX_FRAMES];
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        tx[i] = div_up(fr[i].W, TS_TX);
        ty[i] = div_up(fr[i].H, TS_TY);
    }
    b.map = make_tilemap(n, tx, ty);
    b.win_curr = win_curr;
    b.win_prev = win_prev;
    b.fuse = fuse_erosion ? 1 : 0;
    if (b.map.total == 0) return hipSuccess;
    if (win_curr / 2 == TS_R && win_prev / 2 == 0)
        hipLaunchKernelGGL(screening_tile, dim3(b.map.total), dim3(256), 0, st, b);
    else
        hipLaunchKernelGGL(screening_generic, dim3(b.map.total), dim3(256), 0, st, b);
    return hipGetLastError();
}

}  // namespace xdn
