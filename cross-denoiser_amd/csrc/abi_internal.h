// Internal helpers shared by the C-ABI translation units of libxdn.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/xdn.h"
#include "xdn_cnn.h"
#include "xdn_common.h"
#include "xdn_unet.h"

namespace xdn_abi {

int fail(int code, const char *fmt, ...);  // sets the thread's last-error message, returns code

inline int hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) return fail(XDN_EHIP, "%s: %s", what, hipGetErrorString(e));
    return XDN_OK;
}

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline size_t px(int H, int W) { return (size_t)H * (size_t)W; }

inline int check_shape(int B, int H, int W)
{
    if (B < 0 || H <= 0 || W <= 0) return fail(XDN_EINVAL, "invalid shape [%d, %d, %d]", B, H, W);
    return XDN_OK;
}

struct LayerSpec {
    int ci, co;
};
constexpr LayerSpec kCnn5[5] = {{15, 32}, {32, 32}, {32, 32}, {32, 32}, {32, 6}};

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// The cnn5 kernels address their 64 B/px activations with 32-bit buffer offsets, so a frame has
// fewer than 2^25 pixels (7680x4320 fits); the filter kernels index pixels with 32 bits.
constexpr size_t kCnn5MaxPx = (size_t)1 << 25;
constexpr size_t kFilterMaxPx = (size_t)1 << 32;

// FP32 parameter offsets (floats) and packed-fragment byte offsets of cnn5.
struct Cnn5Layout {
    size_t w_off[5], b_off[5], count;
    size_t fwd_off[5], dg_off[5], bias_off[5], packed;
    Cnn5Layout()
    {
        size_t o = 0, q = 0;
        for (int l = 0; l < 5; ++l)
        {
            w_off[l] = o;
            o += 9 * (size_t)kCnn5[l].ci * kCnn5[l].co;
            b_off[l] = o;
            o += kCnn5[l].co;
            const int kcf = (kCnn5[l].ci + 15) / 16, kcd = (kCnn5[l].co + 15) / 16;
            fwd_off[l] = q;
            q += al256((size_t)9 * kcf * 64 * 16);
            dg_off[l] = q;
            q += l > 0 ? al256((size_t)9 * kcd * 64 * 16) : 0;
            bias_off[l] = q;
            q += 256;
        }
        count = o;
        packed = q;
    }
};
inline const Cnn5Layout kLay;

// Workspace of cnn5 for B images of H x W: hidden activations, gradient ping-pong, and per layer
// the weight-gradient partials of every block and the reduction stage (all five layers are reduced
// together after the backward pass). The backward takes each ReLU derivative from the activation
// itself (x > 0), so no mask is stored.
struct Cnn5Ws {
    uint16_t *X[4], *dY[2];
    float *part[5], *stage[5];
    int nblk[5];     // weight-gradient blocks (partials) per image and layer
    int pstride[5];  // floats per partial (the first and last layers' partials are compact)
    size_t bytes;
    Cnn5Ws(void *base, int B, int H, int W)
    {
        const size_t act = al256((size_t)B * H * W * 32 * 2);
        char *p = static_cast<char *>(base);
        size_t o = 0;
        for (int i = 0; i < 4; ++i) X[i] = reinterpret_cast<uint16_t *>(p + o), o += act;
        for (int i = 0; i < 2; ++i) dY[i] = reinterpret_cast<uint16_t *>(p + o), o += act;
        for (int l = 0; l < 5; ++l)
        {
            // every weight gradient comes from a fused backward kernel on the same tile grid (layer
            // 1's from layer 2's backward)
            nblk[l] = xdn::bwd_blocks(H, W);
            pstride[l] = l == 0 ? xdn::WG_PART_L1 : (l == 4 ? xdn::WG_PART_L5 : xdn::WG_PART);
            part[l] = reinterpret_cast<float *>(p + o);
            o += al256((size_t)B * nblk[l] * pstride[l] * 4);
            stage[l] = reinterpret_cast<float *>(p + o);
            o += al256((size_t)xdn::WG_SPLITS * pstride[l] * 4);
        }
        bytes = o;
    }
};

// Reduction of layer l's nblk_total partials (the B images' of one workspace, or one slot's) into
// its flat gradient at grad + the layer's weight offset.
inline xdn::RedJob cnn5_red_job(const Cnn5Ws &ws, int l, float *grad, int nblk_total)
{
    xdn::RedJob r{ws.part[l], ws.stage[l], grad + kLay.w_off[l], nblk_total, kCnn5[l].ci, kCnn5[l].co};
    if (l == 0 || l == 4)
    {
        r.compact = 1;
        r.ci_cnt = kCnn5[l].ci;
        r.pstride = ws.pstride[l];
    }
    return r;
}

// Layers 2..5 of a cnn5 forward for the image at pixel offset `po` of the workspace's buffers.
inline xdn::FwdStreamFrame cnn5_stream_frame(const Cnn5Ws &ws, const char *packed, size_t po, float *out, int H, int W)
{
    xdn::FwdStreamFrame f{};
    f.x1 = ws.X[0] + po * 32;
    for (int l = 1; l < 4; ++l) f.x[l - 1] = ws.X[l] + po * 32;
    f.out = out;
    for (int l = 1; l < 5; ++l)
    {
        f.wfrag[l - 1] = packed + kLay.fwd_off[l];
        f.bias[l - 1] = reinterpret_cast<const float *>(packed + kLay.bias_off[l]);
    }
    f.H = H;
    f.W = W;
    return f;
}

// ---------------------------------------------------------------------------------------------
// unet: the reference's bandwidth U-Net (scripts/main.py:83-118). Parameters in Keras order, per
// layer kernel then bias: 3x3 kernels [3][3][ci][co], transposed 1x1 kernels [co][ci], the head
// [ci][co]. Levels: 0 = full resolution, 1 = half, 2 = quarter.
// ---------------------------------------------------------------------------------------------
enum { UL_CONV = 0, UL_CONVT = 1, UL_HEAD = 2 };
struct UnetLayer {
    int type, ci, co, level;  // level of the layer's input
};
constexpr int UNET_LAYERS = 12;
constexpr UnetLayer kUnet[UNET_LAYERS] = {
    {UL_CONV, 15, 8, 0},  {UL_CONV, 8, 8, 0},   {UL_CONV, 8, 16, 1},   {UL_CONV, 16, 16, 1},
    {UL_CONV, 16, 32, 2}, {UL_CONV, 32, 32, 2}, {UL_CONVT, 32, 32, 2}, {UL_CONV, 48, 16, 1},
    {UL_CONV, 16, 16, 1}, {UL_CONVT, 16, 16, 1}, {UL_CONV, 24, 8, 0},  {UL_HEAD, 8, 6, 0}};

// FP32 parameter offsets (floats); packed bytes: bf16 fragments, padded biases of the 3x3 layers
// and an FP32 copy of the 1x1 layers' kernel + bias.
struct UnetLayout {
    size_t w_off[UNET_LAYERS], b_off[UNET_LAYERS], count;
    size_t fwd_off[UNET_LAYERS], dg_off[UNET_LAYERS], bias_off[UNET_LAYERS], master_off[UNET_LAYERS], packed;
    UnetLayout()
    {
        size_t o = 0, q = 0;
        for (int l = 0; l < UNET_LAYERS; ++l)
        {
            const UnetLayer &L = kUnet[l];
            const size_t nw = (L.type == UL_CONV ? 9 : 1) * (size_t)L.ci * L.co;
            w_off[l] = o;
            o += nw;
            b_off[l] = o;
            o += L.co;
            fwd_off[l] = dg_off[l] = bias_off[l] = master_off[l] = 0;
            if (L.type == UL_CONV)
            {
                const size_t kcf = (L.ci + 15) / 16, ntf = (L.co + 31) / 32;
                const size_t kcd = (L.co + 15) / 16, ntd = (L.ci + 31) / 32;
                fwd_off[l] = q;
                q += al256(9 * kcf * ntf * 64 * 16);
                dg_off[l] = q;
                q += l > 0 ? al256(9 * kcd * ntd * 64 * 16) : 0;
                bias_off[l] = q;
                q += al256(ntf * 32 * 4);
            }
            else
            {
                master_off[l] = q;
                q += al256((nw + L.co) * 4);
            }
        }
        count = o;
        packed = q;
    }
};
inline const UnetLayout kULay;

// Workspace buffers of the U-Net, [B][h][w][ch] each (bf16 unless f32).
enum {
    UB_A0, UB_A1, UB_P1, UB_A2, UB_A3, UB_P2, UB_A4, UB_A5, UB_A6, UB_A7, UB_A8, UB_A9, UB_A10,
    UB_DY10, UB_DC5, UB_DY8, UB_DY7, UB_DC4, UB_DY5, UB_DY4, UB_DP2, UB_DY3, UB_DY2, UB_DP1, UB_DY1, UB_DY0,
    UNET_BUFS
};
struct UnetBuf {
    int level, ch, f32;
};
constexpr UnetBuf kUnetBuf[UNET_BUFS] = {
    {0, 8, 0},  {0, 8, 0},  {1, 8, 0},  {1, 16, 0}, {1, 16, 0}, {2, 16, 0}, {2, 32, 0},
    {2, 32, 0}, {1, 32, 0}, {1, 16, 0}, {1, 16, 0}, {0, 16, 0}, {0, 8, 0},              // A0 .. A10
    {0, 8, 0},  {0, 24, 0}, {1, 16, 0}, {1, 16, 0}, {1, 48, 0}, {2, 32, 0}, {2, 32, 0},
    {2, 16, 0}, {1, 16, 0}, {1, 16, 0}, {1, 8, 0},  {0, 8, 0},  {0, 8, 0}};             // gradients
// Weight-gradient partial sets: one per 3x3 layer (two for c4_1: input channels 0-31 and 32-47),
// the two transposed layers and the head. level = resolution the set's blocks walk.
constexpr int UNET_JOBS = 13;
struct UnetJob {
    int layer, level;
};
constexpr UnetJob kUnetJob[UNET_JOBS] = {{0, 0}, {1, 0}, {2, 1}, {3, 1}, {4, 2},  {5, 2}, {6, 2},
                                         {7, 1}, {7, 1}, {8, 1}, {9, 1}, {10, 0}, {11, 0}};

// input channels of partial set j's slice (c4_1 is split 32 + 16)
inline int unet_job_slice(int job)
{
    return job == 7 ? 32 : (job == 8 ? 16 : kUnet[kUnetJob[job].layer].ci);
}
// floats of one block's compact partial of set j
inline int unet_job_pstride(int job)
{
    const UnetLayer &L = kUnet[kUnetJob[job].layer];
    return xdn::compact_part(L.type == UL_CONV ? 9 : 1, unet_job_slice(job), L.co);
}

inline int unet_job_blocks(int job, int h, int w)
{
    const int t = kUnet[kUnetJob[job].layer].type;
    const long np = (long)h * w;
    if (t == UL_CONV) return xdn::ublocks(h, w);
    const long per = t == UL_CONVT ? 512 : 4096;  // pixels per block
    const long cap = 512;
    long nb = (np + per - 1) / per;
    nb = nb < 1 ? 1 : (nb > cap ? cap : nb);
    return (int)nb;
}

struct UnetWs {
    char *buf[UNET_BUFS];
    size_t off[UNET_BUFS], frame_bytes[UNET_BUFS];
    float *part[UNET_JOBS], *stage[UNET_JOBS];
    int nblk[UNET_JOBS];
    int h[3], w[3];
    size_t bytes;
    UnetWs(void *base, int B, int H, int W)
    {
        h[0] = H, w[0] = W, h[1] = H / 2, w[1] = W / 2, h[2] = H / 4, w[2] = W / 4;
        char *p = static_cast<char *>(base);
        size_t o = 0;
        for (int i = 0; i < UNET_BUFS; ++i)
        {
            const UnetBuf &U = kUnetBuf[i];
            frame_bytes[i] = (size_t)h[U.level] * w[U.level] * U.ch * (U.f32 ? 4 : 2);
            off[i] = o;
            buf[i] = p + o;
            o += al256(frame_bytes[i] * B);
        }
        for (int j = 0; j < UNET_JOBS; ++j)
        {
            const int lv = kUnetJob[j].level;
            nblk[j] = unet_job_blocks(j, h[lv], w[lv]);
            part[j] = reinterpret_cast<float *>(p + o);
            o += al256((size_t)B * nblk[j] * unet_job_pstride(j) * 4);
            stage[j] = reinterpret_cast<float *>(p + o);
            o += al256((size_t)xdn::WG_SPLITS * unet_job_pstride(j) * 4);
        }
        bytes = o;
    }
    void *at(int i, int b) const { return buf[i] + (size_t)b * frame_bytes[i]; }
};

// One image of a U-Net launch set (abi_unet.cpp): its workspace view, inputs and outputs. The
// images of one set may differ in size and weights; each launch covers all of them.
struct UnetFrame {
    char *buf[UNET_BUFS];
    float *part[UNET_JOBS];
    int nblk[UNET_JOBS], h[3], w[3];
    const float *img[5];   // the five 3-channel network inputs [h0][w0][3]
    const char *pk;        // packed weights
    float *param;          // network output [h0][w0][6]
    const float *dparam;   // its gradient (backward)
};
// image b of a workspace laid out for B images
UnetFrame unet_frame(const UnetWs &ws, int b, const float *const img[5], const void *packed, float *param,
                     const float *dparam);
int unet_forward_frames(const UnetFrame *f, int n, hipStream_t st);
// the backward pass up to the weight-gradient partials (reduced by the caller's jobs)
int unet_backward_frames(const UnetFrame *f, int n, hipStream_t st);
// reduction job of partial set `set` (nblk_total partials from part) into the flat gradient
xdn::RedJob unet_red_job(int set, float *part, float *stage, float *grad, int nblk_total);
// pack jobs of one parameter vector (3x3 fragments and FP32 copies of the 1x1 layers)
void unet_pack_jobs(const float *params, void *packed, std::vector<xdn::UPackJob> &jobs);

}  // namespace xdn_abi
