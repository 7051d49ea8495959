// Backward of cross_regression with respect to its four inputs (imga, imgb, albedo, normal),
// recon_scale 1. The reference registers no gradient for this op (regression.cc:25-103); this is
// the adjoint of the forward kernels of k_regression.hip, with every window weight recomputed.
//
// Notation for one half X (features f = own log colour, v = var(f), target t = other half):
//   regression at center c:  w = exp(-|f_i-f_c|^2/(v_c+v_i+1e-4)) + 1e-7,
//                            d = [1, (f_i-f_c)/(sqrt v_c + sqrt v_i + 1e-2), a_i-a_c, n_i-n_c],
//                            M = sum w d d^T + 1e-3 I,  b = sum w d t_i,  beta = M^-1 b
//   recon at pixel p:        out = max(0, sum_c w_pc (beta_c . delta_pc) / max(sum_c w_pc, 1e-7)),
//                            delta = x_p - x_c.
// Adjoint: u = M^-1 g_beta (the Cholesky clamp never binds: every pivot of M is >= 1e-3), and per
// tap  g_w = (u.d)(t_i - d.beta),  g_d = w (u (t_i - d.beta) - beta (u.d)),  g_t_i = w (u.d).
// Conventions at the kinks: max(0, .) passes the gradient where the value is > 0, the sumW guard
// where sumW > 1e-7, and d sqrt(v)/dv is 0 where v == 0.
//
// Kernels (all batched over frames like the forward; one thread per pixel or per (center, half)):
//   crg_recon_prep   per pixel: recompute the reconstruction sums, emit (g_acc, g_sumW)   [H][W][2]
//   crg_center       per (center, half): gather g_beta over the pixels whose range holds the
//                    center (recon adjoint, center side), rebuild M, solve u, then the
//                    center side of the regression adjoint over the window             [Hs][Ws][2]
//   crg_pixel        per pixel: recon adjoint (pixel side), regression adjoint (tap side,
//                    all mirrored positions of the pixel), center-side terms of grid pixels,
//                    variance adjoint (own term); writes grad albedo/normal and partial
//                    grad imga/imgb
//   crg_var_gather   per pixel: neighbour terms of the variance adjoint
#include "xdn_common.h"
#include "xdn_kernels.h"
#include "xdn_regression.h"

namespace xdn {

struct CrgBatch {
    TileMap map;
    CrgFrame f[XDN_MAX_FRAMES];
    int half_win, sparsity;
};

struct G10 {
    float3 f, a, n;
    float v;
    __device__ __forceinline__ void zero()
    {
        f = a = n = make_float3(0.f, 0.f, 0.f);
        v = 0.f;
    }
};

__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float comp(float3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// Adjoint of (w, delta) for the pair (P, Q), delta = x_P - x_Q, w = exp(-|df|^2 / S) + 1e-7 with
// S = v_P + v_Q + 1e-4, delta_1 = df * inv (inv = 1/(sqrt v_P + sqrt v_Q + 1e-2)).
// Adds the P side (sideP) or the Q side of the gradient to g.
__device__ __forceinline__ void pair_adj(float3 df, float S, float inv, float sP, float sQ, float gw,
                                         const float gd[3][6], bool sideP, G10 &g)
{
    const float D = df.x * df.x + df.y * df.y + df.z * df.z;
    const float rS = fast_rcp(S);
    const float E = fast_exp2(-kLog2e * D * rS);
    const float gD2 = -2.f * gw * E * rS;  // dL/dD * 2
    const float gS = gw * E * D * rS * rS;
    float3 gf, ga;
    gf.x = gD2 * df.x + gd[0][1] * inv;
    gf.y = gD2 * df.y + gd[1][1] * inv;
    gf.z = gD2 * df.z + gd[2][1] * inv;
    ga = make_float3(gd[0][2], gd[1][2], gd[2][2]);
    const float3 gn = make_float3(gd[0][3] + gd[1][3] + gd[2][3], gd[0][4] + gd[1][4] + gd[2][4],
                                  gd[0][5] + gd[1][5] + gd[2][5]);
    const float gden = -(gd[0][1] * df.x + gd[1][1] * df.y + gd[2][1] * df.z) * inv * inv;
    if (sideP)
    {
        g.f.x += gf.x; g.f.y += gf.y; g.f.z += gf.z;
        g.a.x += ga.x; g.a.y += ga.y; g.a.z += ga.z;
        g.n.x += gn.x; g.n.y += gn.y; g.n.z += gn.z;
        g.v += gS + (sP > 0.f ? 0.5f * gden / sP : 0.f);
    }
    else
    {
        g.f.x -= gf.x; g.f.y -= gf.y; g.f.z -= gf.z;
        g.a.x -= ga.x; g.a.y -= ga.y; g.a.z -= ga.z;
        g.n.x -= gn.x; g.n.y -= gn.y; g.n.z -= gn.z;
        g.v += gS + (sQ > 0.f ? 0.5f * gden / sQ : 0.f);
    }
}

// Reconstruction term of pixel p and center c for one half, as recon_generic computes it:
// delta = x_p - x_c; returns w, fills delta d[ch][k] and pred[ch].
__device__ __forceinline__ float recon_pair(float3 fp, float3 ap, float3 np, float vp, float sp, float3 fc, float3 ac,
                                            float3 nc, float vc, float sc, const float *bt, float d[3][6],
                                            float pred[3], float &inv)
{
    const float3 df = sub3(fp, fc), da = sub3(ap, ac), dn = sub3(np, nc);
    const float w = reg_weight(df, vp + 1e-4f, vc);
    inv = fast_rcp((sp + 1e-2f) + sc);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
    {
        d[ch][0] = 1.f;
        d[ch][1] = comp(df, ch) * inv;
        d[ch][2] = comp(da, ch);
        d[ch][3] = dn.x;
        d[ch][4] = dn.y;
        d[ch][5] = dn.z;
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 6; ++k) s += bt[k * 3 + ch] * d[ch][k];
        pred[ch] = s;
    }
    return w;
}

// Regression tap i of center c for one half, as regression_generic computes it: d = x_i - x_c.
__device__ __forceinline__ float reg_pair(float3 fi, float3 ai, float3 ni, float vi, float si, float3 fc, float3 ac,
                                          float3 nc, float vc, float sc, float d[3][6], float &inv)
{
    const float3 df = sub3(fi, fc), da = sub3(ai, ac), dn = sub3(ni, nc);
    const float w = reg_weight(df, vc + 1e-4f, vi);
    inv = fast_rcp((sc + 1e-2f) + si);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
    {
        d[ch][0] = 1.f;
        d[ch][1] = comp(df, ch) * inv;
        d[ch][2] = comp(da, ch);
        d[ch][3] = dn.x;
        d[ch][4] = dn.y;
        d[ch][5] = dn.z;
    }
    return w;
}

// Regression adjoint of one tap: g_w, g_d and the target gradient weight w (u.d) per channel.
__device__ __forceinline__ float reg_tap_adj(float w, const float d[3][6], const float *bt, const float *u, float3 t,
                                             float gd[3][6], float gt[3])
{
    float gw = 0.f;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
    {
        float r = comp(t, ch), s = 0.f;
#pragma unroll
        for (int k = 0; k < 6; ++k)
        {
            r -= d[ch][k] * bt[k * 3 + ch];
            s += u[k * 3 + ch] * d[ch][k];
        }
        gw += s * r;
#pragma unroll
        for (int k = 0; k < 6; ++k) gd[ch][k] = w * (u[k * 3 + ch] * r - bt[k * 3 + ch] * s);
        gt[ch] = w * s;
    }
    return gw;
}

// ---------------------------------------------------------------------------------------------
// 1. per pixel: (g_acc, g_sumW) of the reconstruction, both halves
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void crg_recon_prep(CrgBatch b)
{
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const CrgFrame &F = b.f[L.f];
    const int S = b.sparsity, HW = b.half_win;
    const int H = F.H, W = F.W, Ws = W / S;
    const int px = L.tx * 16 + (threadIdx.x & 15), py = L.ty * 16 + (threadIdx.x >> 4);
    if (px >= W || py >= H) return;
    const size_t p = (size_t)py * W + px;
    const int sx = max(0, idivup(px - HW, S)), ex = min((W - 1) / S, (px + HW) / S);
    const int sy = max(0, idivup(py - HW, S)), ey = min((H - 1) / S, (py + HW) / S);
    const float3 ap = ld3(F.alb, p), np = ld3(F.nrm, p);
    for (int h = 0; h < 2; ++h)
    {
        const float *feat = h ? F.y2 : F.y1, *var = h ? F.var2 : F.var1;
        const float3 fp = ld3(feat, p);
        const float vp = var[p], sp = sqrtf(vp);
        float3 acc = make_float3(0.f, 0.f, 0.f);
        float sum = 0.f;
        for (int iy = sy; iy <= ey; ++iy)
            for (int ix = sx; ix <= ex; ++ix)
            {
                const size_t c = (size_t)(iy * S) * W + (size_t)(ix * S);
                const float *bt = F.beta + (((size_t)iy * Ws + ix) * 2 + h) * 18;
                float d[3][6], pred[3], inv;
                const float vc = var[c];
                const float w = recon_pair(fp, ap, np, vp, sp, ld3(feat, c), ld3(F.alb, c), ld3(F.nrm, c), vc,
                                           sqrtf(vc), bt, d, pred, inv);
                acc.x += w * pred[0];
                acc.y += w * pred[1];
                acc.z += w * pred[2];
                sum += w;
            }
        const float iv = 1.f / fmaxf(sum, 1e-7f);
        const float3 go = ld3(h ? F.goutb : F.gouta, p);
        float4 q;
        q.x = acc.x * iv > 0.f ? go.x * iv : 0.f;
        q.y = acc.y * iv > 0.f ? go.y * iv : 0.f;
        q.z = acc.z * iv > 0.f ? go.z * iv : 0.f;
        q.w = sum > 1e-7f ? -(q.x * acc.x + q.y * acc.y + q.z * acc.z) * iv : 0.f;
        F.gq[p * 2 + h] = q;
    }
}

// ---------------------------------------------------------------------------------------------
// 2. per (center, half): g_beta, u = M^-1 g_beta, center-side gradients
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void crg_center(CrgBatch b)
{
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const CrgFrame &F = b.f[L.f];
    const int S = b.sparsity, HW = b.half_win;
    const int H = F.H, W = F.W, Ws = W / S, Hs = H / S;
    const int h = threadIdx.x >> 7;
    const int bx = L.tx * 16 + (threadIdx.x & 15);
    const int by = L.ty * 8 + ((threadIdx.x & 127) >> 4);
    if (bx >= Ws || by >= Hs) return;
    const float *feat = h ? F.y2 : F.y1, *tgt = h ? F.y1 : F.y2, *var = h ? F.var2 : F.var1;
    const int cx = bx * S, cy = by * S;
    const size_t c = (size_t)cy * W + cx;
    const float3 fc = ld3(feat, c), ac = ld3(F.alb, c), nc = ld3(F.nrm, c);
    const float vc = var[c], sc = sqrtf(vc);
    const size_t crec = ((size_t)by * Ws + bx) * 2 + h;
    float bt[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) bt[k] = F.beta[crec * 18 + k];
    G10 g;
    g.zero();
    float gb[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) gb[k] = 0.f;

    // (a) reconstruction adjoint, gathered over the pixels whose center range holds (by, bx)
    const int py0 = max(0, cy - HW - S), py1 = min(H - 1, cy + HW + S);
    const int px0 = max(0, cx - HW - S), px1 = min(W - 1, cx + HW + S);
    for (int py = py0; py <= py1; ++py)
    {
        if (by < max(0, idivup(py - HW, S)) || by > min((H - 1) / S, (py + HW) / S)) continue;
        for (int px = px0; px <= px1; ++px)
        {
            if (bx < max(0, idivup(px - HW, S)) || bx > min((W - 1) / S, (px + HW) / S)) continue;
            const size_t p = (size_t)py * W + px;
            const float4 q = F.gq[p * 2 + h];
            if (q.x == 0.f && q.y == 0.f && q.z == 0.f && q.w == 0.f) continue;
            const float3 fp = ld3(feat, p);
            const float vp = var[p], sp = sqrtf(vp);
            float d[3][6], pred[3], inv, gd[3][6];
            const float w = recon_pair(fp, ld3(F.alb, p), ld3(F.nrm, p), vp, sp, fc, ac, nc, vc, sc, bt, d, pred, inv);
            const float ga[3] = {q.x, q.y, q.z};
            float gw = q.w;
#pragma unroll
            for (int ch = 0; ch < 3; ++ch)
            {
                gw += ga[ch] * pred[ch];
                const float wg = w * ga[ch];
#pragma unroll
                for (int k = 0; k < 6; ++k)
                {
                    gb[k * 3 + ch] += wg * d[ch][k];
                    gd[ch][k] = wg * bt[k * 3 + ch];
                }
            }
            pair_adj(sub3(fp, fc), (vp + 1e-4f) + vc, inv, sp, sc, gw, gd, false, g);
        }
    }

    // (b) normal equations of the center (as the forward) and u = M^-1 g_beta per channel
    RegAcc acc;
    acc.zero();
    const float cvar_e = vc + 1e-4f, cstd_e = sc + 1e-2f;
    for (int iy = cy - HW; iy <= cy + HW; ++iy)
        for (int ix = cx - HW; ix <= cx + HW; ++ix)
        {
            const size_t i = (size_t)mirror(iy, H) * W + mirror(ix, W);
            const float3 iw = ld3(feat, i), ia = ld3(F.alb, i), in = ld3(F.nrm, i), it = ld3(tgt, i);
            const float ivar = var[i];
            const float3 dw = sub3(iw, fc);
            const float w = reg_weight(dw, cvar_e, ivar);
            const float inv = fast_rcp(cstd_e + sqrtf(ivar));
            acc.add(w, inv, dw, sub3(ia, ac), sub3(in, nc), it);
        }
    float u[18];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
    {
        Sys6 s;
        acc.system(ch, s);
#pragma unroll
        for (int k = 0; k < 6; ++k) s.x[k] = gb[k * 3 + ch];
        float uc[6];
        solve6(s, uc);
#pragma unroll
        for (int k = 0; k < 6; ++k) u[k * 3 + ch] = uc[k];
    }

    // (c) regression adjoint, center side
    for (int iy = cy - HW; iy <= cy + HW; ++iy)
        for (int ix = cx - HW; ix <= cx + HW; ++ix)
        {
            const size_t i = (size_t)mirror(iy, H) * W + mirror(ix, W);
            const float3 fi = ld3(feat, i);
            const float vi = var[i], si = sqrtf(vi);
            float d[3][6], inv, gd[3][6], gt[3];
            const float w = reg_pair(fi, ld3(F.alb, i), ld3(F.nrm, i), vi, si, fc, ac, nc, vc, sc, d, inv);
            const float gw = reg_tap_adj(w, d, bt, u, ld3(tgt, i), gd, gt);
            pair_adj(sub3(fi, fc), (vc + 1e-4f) + vi, inv, si, sc, gw, gd, false, g);
        }

    float *uo = F.u + crec * 18;
#pragma unroll
    for (int k = 0; k < 18; ++k) uo[k] = u[k];
    float4 *go = reinterpret_cast<float4 *>(F.gcen + crec * 12);
    go[0] = make_float4(g.f.x, g.f.y, g.f.z, g.a.x);
    go[1] = make_float4(g.a.y, g.a.z, g.n.x, g.n.y);
    go[2] = make_float4(g.n.z, g.v, 0.f, 0.f);
}

// ---------------------------------------------------------------------------------------------
// 3. per pixel: pixel-side and tap-side adjoints, both halves
// ---------------------------------------------------------------------------------------------
// Tap positions j (along one axis of size n) that the reflect-101 index maps onto pixel i.
__device__ __forceinline__ int mirror_sources(int i, int n, int js[3])
{
    int m = 0;
    js[m++] = i;
    if (i >= 1) js[m++] = -i;
    if (i <= n - 2) js[m++] = 2 * (n - 1) - i;
    return m;
}

__device__ __forceinline__ int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

__global__ __launch_bounds__(256) void crg_pixel(CrgBatch b)
{
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const CrgFrame &F = b.f[L.f];
    const int S = b.sparsity, HW = b.half_win;
    const int H = F.H, W = F.W, Ws = W / S, Hs = H / S;
    const int px = L.tx * 16 + (threadIdx.x & 15), py = L.ty * 16 + (threadIdx.x >> 4);
    if (px >= W || py >= H) return;
    const size_t p = (size_t)py * W + px;
    const float3 ap = ld3(F.alb, p), np = ld3(F.nrm, p);
    const int sx = max(0, idivup(px - HW, S)), ex = min((W - 1) / S, (px + HW) / S);
    const int sy = max(0, idivup(py - HW, S)), ey = min((H - 1) / S, (py + HW) / S);
    int jys[3], jxs[3];
    const int njy = mirror_sources(py, H, jys), njx = mirror_sources(px, W, jxs);
    const bool grid = (px % S == 0) && (py % S == 0);
    // 3x3 clamped neighbourhood (variance adjoint)
    const int nx0 = max(0, px - 1), nx1 = min(W - 1, px + 1), ny0 = max(0, py - 1), ny1 = min(H - 1, py + 1);
    const float ncount = (float)((nx1 - nx0 + 1) * (ny1 - ny0 + 1) - 1);

    float3 gy_own[2], gy_oth[2];
    float3 galb = make_float3(0.f, 0.f, 0.f), gnrm = galb;
    float3 ev[2];
    for (int h = 0; h < 2; ++h)
    {
        const float *feat = h ? F.y2 : F.y1, *tgt = h ? F.y1 : F.y2, *var = h ? F.var2 : F.var1;
        const float3 fp = ld3(feat, p), tp = ld3(tgt, p);
        const float vp = var[p], sp = sqrtf(vp);
        G10 g;
        g.zero();
        float3 gt = make_float3(0.f, 0.f, 0.f);

        // (a) reconstruction adjoint, pixel side
        const float4 q = F.gq[p * 2 + h];
        if (!(q.x == 0.f && q.y == 0.f && q.z == 0.f && q.w == 0.f))
        {
            const float ga[3] = {q.x, q.y, q.z};
            for (int iy = sy; iy <= ey; ++iy)
                for (int ix = sx; ix <= ex; ++ix)
                {
                    const size_t c = (size_t)(iy * S) * W + (size_t)(ix * S);
                    const float *bt = F.beta + (((size_t)iy * Ws + ix) * 2 + h) * 18;
                    const float3 fc = ld3(feat, c);
                    const float vc = var[c], sc = sqrtf(vc);
                    float d[3][6], pred[3], inv, gd[3][6];
                    const float w = recon_pair(fp, ap, np, vp, sp, fc, ld3(F.alb, c), ld3(F.nrm, c), vc, sc, bt, d,
                                               pred, inv);
                    float gw = q.w;
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch)
                    {
                        gw += ga[ch] * pred[ch];
#pragma unroll
                        for (int k = 0; k < 6; ++k) gd[ch][k] = w * ga[ch] * bt[k * 3 + ch];
                    }
                    pair_adj(sub3(fp, fc), (vp + 1e-4f) + vc, inv, sp, sc, gw, gd, true, g);
                }
        }

        // (b) regression adjoint, tap side: every (center, tap position) whose tap maps onto p
        for (int a = 0; a < njy; ++a)
        {
            const int jy = jys[a];
            const int by0 = max(0, floor_div(jy - HW + S - 1, S)), by1 = min(Hs - 1, floor_div(jy + HW, S));
            for (int by = by0; by <= by1; ++by)
                for (int e = 0; e < njx; ++e)
                {
                    const int jx = jxs[e];
                    const int bx0 = max(0, floor_div(jx - HW + S - 1, S)), bx1 = min(Ws - 1, floor_div(jx + HW, S));
                    for (int bx = bx0; bx <= bx1; ++bx)
                    {
                        const size_t c = (size_t)(by * S) * W + (size_t)(bx * S);
                        const size_t crec = ((size_t)by * Ws + bx) * 2 + h;
                        const float *bt = F.beta + crec * 18;
                        const float *u = F.u + crec * 18;
                        const float3 fc = ld3(feat, c);
                        const float vc = var[c], sc = sqrtf(vc);
                        float d[3][6], inv, gd[3][6], gtt[3];
                        const float w = reg_pair(fp, ap, np, vp, sp, fc, ld3(F.alb, c), ld3(F.nrm, c), vc, sc, d, inv);
                        const float gw = reg_tap_adj(w, d, bt, u, tp, gd, gtt);
                        gt.x += gtt[0];
                        gt.y += gtt[1];
                        gt.z += gtt[2];
                        pair_adj(sub3(fp, fc), (vc + 1e-4f) + vp, inv, sp, sc, gw, gd, true, g);
                    }
                }
        }

        // (c) center-side terms when p is a regression center
        if (grid)
        {
            const float4 *gc =
                reinterpret_cast<const float4 *>(F.gcen + (((size_t)(py / S) * Ws + px / S) * 2 + h) * 12);
            const float4 g0 = gc[0], g1 = gc[1], g2 = gc[2];
            g.f.x += g0.x; g.f.y += g0.y; g.f.z += g0.z;
            g.a.x += g0.w; g.a.y += g1.x; g.a.z += g1.y;
            g.n.x += g1.z; g.n.y += g1.w; g.n.z += g2.x;
            g.v += g2.y;
        }

        // (d) variance adjoint, own term: e = g_v * 2 (mean_neighbours - f_p)
        const float *gvx = h ? F.gvarb : F.gvara;
        const float gv = g.v + (gvx ? gvx[p] : 0.f);
        float3 m = make_float3(0.f, 0.f, 0.f);
        for (int yy = ny0; yy <= ny1; ++yy)
            for (int xx = nx0; xx <= nx1; ++xx)
            {
                if (yy == py && xx == px) continue;
                const float3 f = ld3(feat, (size_t)yy * W + xx);
                m.x += f.x; m.y += f.y; m.z += f.z;
            }
        const float in = 1.f / ncount;
        const float3 e = make_float3(2.f * gv * (m.x * in - fp.x), 2.f * gv * (m.y * in - fp.y),
                                     2.f * gv * (m.z * in - fp.z));
        gy_own[h] = sub3(g.f, e);
        gy_oth[h] = gt;
        ev[h] = make_float3(e.x * in, e.y * in, e.z * in);
        galb.x += g.a.x; galb.y += g.a.y; galb.z += g.a.z;
        gnrm.x += g.n.x; gnrm.y += g.n.y; gnrm.z += g.n.z;
    }
    // half A: features imga, target imgb; half B: features imgb, target imga
    st3(F.gya, p, make_float3(gy_own[0].x + gy_oth[1].x, gy_own[0].y + gy_oth[1].y, gy_own[0].z + gy_oth[1].z));
    st3(F.gyb, p, make_float3(gy_own[1].x + gy_oth[0].x, gy_own[1].y + gy_oth[0].y, gy_own[1].z + gy_oth[0].z));
    st3(F.galb, p, galb);
    st3(F.gnrm, p, gnrm);
    float *eo = F.evar + p * 6;
    eo[0] = ev[0].x; eo[1] = ev[0].y; eo[2] = ev[0].z;
    eo[3] = ev[1].x; eo[4] = ev[1].y; eo[5] = ev[1].z;
}

// ---------------------------------------------------------------------------------------------
// 4. per pixel: neighbour terms of the variance adjoint (the 3x3 relation is symmetric)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void crg_var_gather(CrgBatch b)
{
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const CrgFrame &F = b.f[L.f];
    const int H = F.H, W = F.W;
    const int px = L.tx * 16 + (threadIdx.x & 15), py = L.ty * 16 + (threadIdx.x >> 4);
    if (px >= W || py >= H) return;
    const size_t p = (size_t)py * W + px;
    float sa[3] = {0.f, 0.f, 0.f}, sb[3] = {0.f, 0.f, 0.f};
    for (int yy = max(0, py - 1); yy <= min(H - 1, py + 1); ++yy)
        for (int xx = max(0, px - 1); xx <= min(W - 1, px + 1); ++xx)
        {
            if (yy == py && xx == px) continue;
            const float *e = F.evar + ((size_t)yy * W + xx) * 6;
            sa[0] += e[0]; sa[1] += e[1]; sa[2] += e[2];
            sb[0] += e[3]; sb[1] += e[4]; sb[2] += e[5];
        }
    const float3 ga = ld3(F.gya, p), gb = ld3(F.gyb, p);
    st3(F.gya, p, make_float3(ga.x + sa[0], ga.y + sa[1], ga.z + sa[2]));
    st3(F.gyb, p, make_float3(gb.x + sb[0], gb.y + sb[1], gb.z + sb[2]));
}

// ---------------------------------------------------------------------------------------------
static int div_up_g(int a, int b) { return (a + b - 1) / b; }

hipError_t launch_cross_regression_grad(int n, const CrgFrame *fr, int win, int s, hipStream_t st)
{
    CrgBatch b{};
    int tx[XDN_MAX_FRAMES], ty[XDN_MAX_FRAMES], cx[XDN_MAX_FRAMES], cy[XDN_MAX_FRAMES];
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        tx[i] = div_up_g(fr[i].W, 16);
        ty[i] = div_up_g(fr[i].H, 16);
        cx[i] = div_up_g(fr[i].W / s, 16);
        cy[i] = div_up_g(fr[i].H / s, 8);
    }
    b.half_win = win / 2;
    b.sparsity = s;
    b.map = make_tilemap(n, tx, ty);
    if (b.map.total == 0) return hipSuccess;
    hipLaunchKernelGGL(crg_recon_prep, dim3(b.map.total), dim3(256), 0, st, b);
    CrgBatch bc = b;
    bc.map = make_tilemap(n, cx, cy);
    if (bc.map.total > 0) hipLaunchKernelGGL(crg_center, dim3(bc.map.total), dim3(256), 0, st, bc);
    hipLaunchKernelGGL(crg_pixel, dim3(b.map.total), dim3(256), 0, st, b);
    hipLaunchKernelGGL(crg_var_gather, dim3(b.map.total), dim3(256), 0, st, b);
    return hipGetLastError();
}

}  // namespace xdn
