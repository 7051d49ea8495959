// Motion-vector reprojection on gfx950 (Reproject / isReprojValid, reproject.cu.cc:32-238).
//
// One lane per pixel, 64x4 pixels per workgroup (full-wave rows for coalesced NHWC loads).
// HBM-bound gather. The validity mask and the tap indices must equal the reference bit for bit,
// so the decision path has a pinned FP32 operation order (no contraction in this file):
//   posPrev   = fmaf(mvec, (float)dim, (float)pos)                      (:97, :103)
//   iposPrev  = trunc(posPrev + 0.5f), tap base = trunc(posPrev)          (cuda_utils.cuh:255-258)
//   depth     = |Zp - Z| / (fwZ + 1e-2f) > 10                             correctly rounded FP32 div
//   normal    = sqrtf(fmaf(dz,dz, fmaf(dx,dx, dy*dy))) / ((double)fwN + 1e-2) > 4   (FP64 divide)
//   bilinear  = fractions posPrev - floor(posPrev), sum of the valid weights in tap order >= 0.01f
// The buffer list is passed by value (no pointer table in device memory, no host sync).
#include "xdn_common.h"
#include "xdn_kernels.h"

#pragma clang fp contract(off)

namespace xdn {

struct RpBatch {
    TileMap map;
    int n_list;
    int dims[XDN_MAX_LIST];
    RpFrame f[XDN_MAX_FRAMES];
};
static_assert(sizeof(RpBatch) <= 4096, "kernel argument too large");

constexpr int RP_TX = 64, RP_TY = 4;

__device__ __forceinline__ bool reproj_valid(int W, int H, int px, int py, float Z, float Zp, float fwZ, float3 n,
                                             float3 np, float fwN)
{
    if (px < 1 || py < 1 || px > W - 1 || py > H - 1) return false;
    if (fabsf(Zp - Z) / (fwZ + 1e-2f) > 10.f) return false;
    const float dx = n.x - np.x, dy = n.y - np.y, dz = n.z - np.z;
    const float d = __builtin_sqrtf(__builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy)));
    if ((double)d / ((double)fwN + 1e-2) > 4.0) return false;
    return true;
}

// FNL > 0: the buffer list is FNL buffers of FD channels, known at compile time (the training
// step reprojects three 3-channel buffers); FNL == 0 reads the list shape from the batch.
// UNIT_OPAC: both opacities are one (the training step's case): their loads are skipped.
template <int FNL, int FD, bool UNIT_OPAC>
__global__ __launch_bounds__(256) void reproject_kernel(RpBatch b)
{
    const int tile = xcd_tile(blockIdx.x, b.map.total);
    const TileLoc L = locate(b.map, tile);
    const RpFrame &F = b.f[L.f];
    const int H = F.H, W = F.W;
    const int cx = L.tx * RP_TX + (threadIdx.x & (RP_TX - 1)), cy = L.ty * RP_TY + threadIdx.x / RP_TX;
    if (cx >= W || cy >= H) return;
    const size_t c = (size_t)cy * W + cx;
    const int nl = FNL > 0 ? FNL : b.n_list;
    float acc[XDN_MAX_LIST][4];
#pragma unroll
    for (int l = 0; l < XDN_MAX_LIST; ++l)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[l][k] = 0.f;

    // Three round trips to memory: the pixel's own inputs; every bilinear tap's validity inputs,
    // loaded at coordinates clamped into the image and used only where the tap is in the image;
    // the list values of all four taps (compile-time lists; only the valid taps are summed). The
    // 3x3 fallback of a pixel with no valid tap visits its neighbours one at a time: loading them
    // all at once needs 120 registers per lane and was slower.
    const float op = UNIT_OPAC ? 1.f : F.opac[c];
    const float mx = F.mvec[3 * c], my = F.mvec[3 * c + 1];
    const float fwN = F.pnFw[3 * c + 1];
    const float Z = F.linZ[3 * c], fwZ = F.linZ[3 * c + 1];
    const float3 n = ld3(F.nrm, c);
    bool valid = false;
    if (op != 0.f)
    {
        const float ppx = __builtin_fmaf(mx, (float)W, (float)cx);
        const float ppy = __builtin_fmaf(my, (float)H, (float)cy);
        const int ipx = (int)(ppx + 0.5f), ipy = (int)(ppy + 0.5f);
        const int bx = (int)ppx, by = (int)ppy;
        bool inb[4];
        size_t li[4];
        float po[4], pz[4];
        float3 pn[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
        {
            const int lx = bx + (t & 1), ly = by + (t >> 1);
            inb[t] = !(lx >= W || lx < 0 || ly >= H || ly < 0);
            li[t] = (size_t)min(max(ly, 0), H - 1) * W + min(max(lx, 0), W - 1);
            po[t] = UNIT_OPAC ? 1.f : F.prevOpac[li[t]];
            pz[t] = F.prevLinZ[3 * li[t]];
            pn[t] = ld3(F.prevNrm, li[t]);
        }
        bool v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
        {
            v[t] = inb[t] && po[t] != 0.f && reproj_valid(W, H, ipx, ipy, Z, pz[t], fwZ, n, pn[t], fwN);
            valid = valid || v[t];
        }
        if (valid && FNL > 0)
        {
            // every tap's list values in flight at once (clamped in-image indices, so the loads are
            // in bounds), then summed in tap order where the tap is valid: the same sums as below
            const float x = ppx - floorf(ppx), y = ppy - floorf(ppy);
            const float w[4] = {(1.f - x) * (1.f - y), x * (1.f - y), (1.f - x) * y, x * y};
            float val[4][FNL > 0 ? FNL : 1][FNL > 0 ? FD : 1];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int l = 0; l < FNL; ++l)
#pragma unroll
                    for (int k = 0; k < FD; ++k) val[t][l][k] = F.in[l][li[t] * FD + k];
            float sumw = 0.f;
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
#pragma unroll
                for (int l = 0; l < FNL; ++l)
#pragma unroll
                    for (int k = 0; k < FD; ++k) acc[l][k] = v[t] ? __builtin_fmaf(w[t], val[t][l][k], acc[l][k]) : acc[l][k];
                sumw = v[t] ? sumw + w[t] : sumw;
            }
            valid = (sumw >= 0.01f);
            const float inv = 1.f / sumw;
#pragma unroll
            for (int l = 0; l < XDN_MAX_LIST; ++l)
#pragma unroll
                for (int k = 0; k < 4; ++k) acc[l][k] = valid ? acc[l][k] * inv : 0.f;
        }
        else if (valid)
        {
            const float x = ppx - floorf(ppx), y = ppy - floorf(ppy);
            const float w[4] = {(1.f - x) * (1.f - y), x * (1.f - y), (1.f - x) * y, x * y};
            float sumw = 0.f;
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                if (!v[t]) continue;
#pragma unroll
                for (int l = 0; l < XDN_MAX_LIST; ++l)
                {
                    if (l >= nl) break;
                    const int d = FNL > 0 ? FD : b.dims[l];
                    const float *src = F.in[l] + li[t] * d;
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (k < d) acc[l][k] = __builtin_fmaf(w[t], src[k], acc[l][k]);
                }
                sumw += w[t];
            }
            valid = (sumw >= 0.01f);
            const float inv = 1.f / sumw;
#pragma unroll
            for (int l = 0; l < XDN_MAX_LIST; ++l)
#pragma unroll
                for (int k = 0; k < 4; ++k) acc[l][k] = valid ? acc[l][k] * inv : 0.f;
        }
        if (!valid)
        {
            int nvalid = 0;
            for (int yy = -1; yy <= 1; ++yy)
                for (int xx = -1; xx <= 1; ++xx)
                {
                    const int qx = ipx + xx, qy = ipy + yy;
                    if (qx >= W || qx < 0 || qy >= H || qy < 0) continue;
                    const size_t qi = (size_t)qy * W + qx;
                    if (!UNIT_OPAC && F.prevOpac[qi] == 0.f) continue;
                    if (!reproj_valid(W, H, ipx, ipy, Z, F.prevLinZ[3 * qi], fwZ, n, ld3(F.prevNrm, qi), fwN)) continue;
#pragma unroll
                    for (int l = 0; l < XDN_MAX_LIST; ++l)
                    {
                        if (l >= nl) break;
                        const int d = FNL > 0 ? FD : b.dims[l];
                        const float *src = F.in[l] + qi * d;
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            if (k < d) acc[l][k] += src[k];
                    }
                    ++nvalid;
                }
            if (nvalid > 0)
            {
                valid = true;
                const float inv = 1.f / (float)nvalid;
#pragma unroll
                for (int l = 0; l < XDN_MAX_LIST; ++l)
#pragma unroll
                    for (int k = 0; k < 4; ++k) acc[l][k] *= inv;
            }
            else
            {
#pragma unroll
                for (int l = 0; l < XDN_MAX_LIST; ++l)
#pragma unroll
                    for (int k = 0; k < 4; ++k) acc[l][k] = 0.f;
            }
        }
    }
#pragma unroll
    for (int l = 0; l < XDN_MAX_LIST; ++l)
    {
        if (l >= nl) break;
        const int d = FNL > 0 ? FD : b.dims[l];
        float *dst = F.out[l] + c * d;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < d) dst[k] = acc[l][k];
    }
    F.success[c] = valid ? 1 : 0;
}

static int div_up(int a, int b) { return (a + b - 1) / b; }

hipError_t launch_reproject(int n, const RpFrame *fr, int n_list, const int *dims, bool unit_opacity,
                            hipStream_t st)
{
    RpBatch b{};
    int tx[XDN_MAX_FRAMES], ty[XDN_MAX_FRAMES];
    b.n_list = n_list;
    for (int l = 0; l < n_list; ++l) b.dims[l] = dims[l];
    for (int i = 0; i < n; ++i)
    {
        b.f[i] = fr[i];
        tx[i] = div_up(fr[i].W, RP_TX);
        ty[i] = div_up(fr[i].H, RP_TY);
    }
    b.map = make_tilemap(n, tx, ty);
    if (b.map.total == 0) return hipSuccess;
    bool three3 = n_list == 3;
    for (int l = 0; l < n_list; ++l) three3 = three3 && dims[l] == 3;
    if (three3 && unit_opacity)
        hipLaunchKernelGGL((reproject_kernel<3, 3, true>), dim3(b.map.total), dim3(256), 0, st, b);
    else if (three3)
        hipLaunchKernelGGL((reproject_kernel<3, 3, false>), dim3(b.map.total), dim3(256), 0, st, b);
    else
        hipLaunchKernelGGL((reproject_kernel<0, 0, false>), dim3(b.map.total), dim3(256), 0, st, b);
    return hipGetLastError();
}

}  // namespace xdn
