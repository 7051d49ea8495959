// The U-Net (the reference's bandwidth network, scripts/main.py:83-118): forward and backward as
// launches of the k_unet.hip kernels over a set of images (each with its own size, workspace and
// weights), and the extern "C" entry points of include/xdn.h on top of them.
#include <hip/hip_runtime.h>
#include <string.h>

#include <vector>

#include "abi_internal.h"
#include "xdn_misc.h"
#include "xdn_unet.h"

using namespace xdn_abi;

namespace {

using Frames = const UnetFrame *;

// 3x3 convolution of layer l over every image: forward from concat[a, b] (the images when a < 0)
// into buffer out, or (dgrad) data gradient from buffer a into out, masked by buffer mask's ReLU
// (unmasked when mask < 0).
int conv(Frames F, int n, hipStream_t st, int kind, int l, int a, int b, int out, int mask, bool dgrad,
         const char *what)
{
    const int lv = kUnet[l].level;
    std::vector<xdn::UcFrame> fr(n);
    for (int i = 0; i < n; ++i)
    {
        const UnetFrame &U = F[i];
        xdn::UcFrame &f = fr[i];
        memset(&f, 0, sizeof(f));
        if (a >= 0)
            f.a = U.buf[a];
        else
            for (int k = 0; k < 5; ++k) f.img[k] = U.img[k];
        f.b = b >= 0 ? U.buf[b] : nullptr;
        f.out = U.buf[out];
        f.mask = mask >= 0 ? reinterpret_cast<const uint16_t *>(U.buf[mask]) : nullptr;
        f.wfrag = U.pk + (dgrad ? kULay.dg_off[l] : kULay.fwd_off[l]);
        f.bias = reinterpret_cast<const float *>(U.pk + kULay.bias_off[l]);
        if (kind == xdn::UC_F_24_8_HEAD)
        {
            f.head_w = reinterpret_cast<const float *>(U.pk + kULay.master_off[11]);
            f.head_out = U.param;
        }
        f.H = U.h[lv];
        f.W = U.w[lv];
    }
    return hip_check(xdn::launch_uconv(kind, n, fr.data(), st), what);
}

int pool(Frames F, int n, hipStream_t st, int C, int in, int out)
{
    const int lv = kUnetBuf[in].level;
    std::vector<xdn::PoolFrame> fr(n);
    for (int i = 0; i < n; ++i)
        fr[i] = xdn::PoolFrame{reinterpret_cast<const uint16_t *>(F[i].buf[in]),
                               reinterpret_cast<uint16_t *>(F[i].buf[out]), F[i].h[lv], F[i].w[lv]};
    return hip_check(xdn::launch_maxpool(C, n, fr.data(), st), "unet max pool");
}

const float *master(const UnetFrame &U, int l) { return reinterpret_cast<const float *>(U.pk + kULay.master_off[l]); }

// transposed layer l forward: buffer in -> buffer out
int convT(Frames F, int n, hipStream_t st, int l, int in, int out)
{
    const int lv = kUnet[l].level;
    std::vector<xdn::CtFrame> fr(n);
    for (int i = 0; i < n; ++i)
    {
        xdn::CtFrame &f = fr[i];
        memset(&f, 0, sizeof(f));
        f.in = reinterpret_cast<const uint16_t *>(F[i].buf[in]);
        f.out = reinterpret_cast<uint16_t *>(F[i].buf[out]);
        f.w = master(F[i], l);
        f.b = f.w + kUnet[l].ci * kUnet[l].co;
        f.h = F[i].h[lv];
        f.w_ = F[i].w[lv];
    }
    return hip_check(xdn::launch_convT(kUnet[l].co, n, fr.data(), st), "unet transposed conv");
}

// weight gradient of partial set j over concat[a, b] (or the images) and dY buffer dy
int wgrad(Frames F, int n, hipStream_t st, int kind, int j, int a, int b, int dy)
{
    const int lv = kUnetJob[j].level;
    std::vector<xdn::UwFrame> fr(n);
    for (int i = 0; i < n; ++i)
    {
        const UnetFrame &U = F[i];
        xdn::UwFrame &f = fr[i];
        memset(&f, 0, sizeof(f));
        if (a >= 0)
            f.a = U.buf[a];
        else
            for (int k = 0; k < 5; ++k) f.img[k] = U.img[k];
        f.b = b >= 0 ? U.buf[b] : nullptr;
        f.dy = reinterpret_cast<const uint16_t *>(U.buf[dy]);
        f.part = U.part[j];
        f.nblk = U.nblk[j];
        f.H = U.h[lv];
        f.W = U.w[lv];
    }
    return hip_check(xdn::launch_uwgrad(kind, n, fr.data(), st), "unet wgrad");
}

// transposed layer l backward (partial set j): input buffer in, gradient of its output in buffer
// dout (its first C channels); the gradient into its input -> din, masked by in's ReLU
int convT_bwd(Frames F, int n, hipStream_t st, int l, int j, int in, int dout, int din)
{
    const int lv = kUnet[l].level;
    std::vector<xdn::CtFrame> fr(n);
    for (int i = 0; i < n; ++i)
    {
        const UnetFrame &U = F[i];
        xdn::CtFrame &f = fr[i];
        memset(&f, 0, sizeof(f));
        f.in = reinterpret_cast<const uint16_t *>(U.buf[in]);
        f.dout = reinterpret_cast<const uint16_t *>(U.buf[dout]);
        f.din = reinterpret_cast<uint16_t *>(U.buf[din]);
        f.w = master(U, l);
        f.b = f.w + kUnet[l].ci * kUnet[l].co;
        f.part = U.part[j];
        f.h = U.h[lv];
        f.w_ = U.w[lv];
        f.dstride = kUnetBuf[dout].ch;
        f.nblk = U.nblk[j];
    }
    return hip_check(xdn::launch_convT_bwd(kUnet[l].co, n, fr.data(), st), "unet transposed conv backward");
}

// gradient into a pooled activation: pooled gradient dp + skip gradient (channels soff.. of
// buffer skip) -> dy, masked by act's ReLU
int pool_bwd(Frames F, int n, hipStream_t st, int C, int dp, int act, int skip, int soff, int dy)
{
    const int lv = kUnetBuf[act].level;
    std::vector<xdn::PoolBwFrame> fr(n);
    for (int i = 0; i < n; ++i)
    {
        const UnetFrame &U = F[i];
        fr[i] = xdn::PoolBwFrame{reinterpret_cast<const uint16_t *>(U.buf[dp]),
                                 reinterpret_cast<const uint16_t *>(U.buf[act]),
                                 reinterpret_cast<const uint16_t *>(U.buf[skip]),
                                 reinterpret_cast<uint16_t *>(U.buf[dy]), U.h[lv], U.w[lv], kUnetBuf[skip].ch, soff};
    }
    return hip_check(xdn::launch_pool_bwd(C, n, fr.data(), st), "unet max pool backward");
}

// head backward: dY10 = ReLU'(A10) * W dparam, partial set 12 (the forward runs in c5_1's epilogue)
int head_bwd(Frames F, int n, hipStream_t st)
{
    std::vector<xdn::HeadFrame> fr(n);
    for (int i = 0; i < n; ++i)
    {
        const UnetFrame &U = F[i];
        xdn::HeadFrame &f = fr[i];
        memset(&f, 0, sizeof(f));
        f.in = reinterpret_cast<const uint16_t *>(U.buf[UB_A10]);
        f.dparam = U.dparam;
        f.din = reinterpret_cast<uint16_t *>(U.buf[UB_DY10]);
        f.w = master(U, 11);
        f.b = f.w + 48;
        f.part = U.part[12];
        f.H = U.h[0];
        f.W = U.w[0];
        f.nblk = U.nblk[12];
    }
    return hip_check(xdn::launch_head_bwd(n, fr.data(), st), "unet head backward");
}

}  // namespace

namespace xdn_abi {

UnetFrame unet_frame(const UnetWs &ws, int b, const float *const img[5], const void *packed, float *param,
                     const float *dparam)
{
    UnetFrame U;
    for (int i = 0; i < UNET_BUFS; ++i) U.buf[i] = static_cast<char *>(ws.at(i, b));
    for (int j = 0; j < UNET_JOBS; ++j)
    {
        U.part[j] = ws.part[j] + (size_t)b * ws.nblk[j] * unet_job_pstride(j);
        U.nblk[j] = ws.nblk[j];
    }
    for (int k = 0; k < 3; ++k) U.h[k] = ws.h[k], U.w[k] = ws.w[k];
    for (int k = 0; k < 5; ++k) U.img[k] = img[k];
    U.pk = static_cast<const char *>(packed);
    U.param = param;
    U.dparam = dparam;
    return U;
}

int unet_forward_frames(const UnetFrame *F, int n, hipStream_t st)
{
    using namespace xdn;
    int e = 0;
    if ((e = conv(F, n, st, UC_F_0, 0, -1, -1, UB_A0, -1, false, "unet c1_1"))) return e;
    if ((e = conv(F, n, st, UC_F_8_8, 1, UB_A0, -1, UB_A1, -1, false, "unet c1_2"))) return e;
    if ((e = pool(F, n, st, 8, UB_A1, UB_P1))) return e;
    if ((e = conv(F, n, st, UC_F_8_16, 2, UB_P1, -1, UB_A2, -1, false, "unet c2_1"))) return e;
    if ((e = conv(F, n, st, UC_F_16_16, 3, UB_A2, -1, UB_A3, -1, false, "unet c2_2"))) return e;
    if ((e = pool(F, n, st, 16, UB_A3, UB_P2))) return e;
    if ((e = conv(F, n, st, UC_F_16_32, 4, UB_P2, -1, UB_A4, -1, false, "unet c3_1"))) return e;
    if ((e = conv(F, n, st, UC_F_32_32, 5, UB_A4, -1, UB_A5, -1, false, "unet c3_2"))) return e;
    if ((e = convT(F, n, st, 6, UB_A5, UB_A6))) return e;                                       // c3_3
    if ((e = conv(F, n, st, UC_F_48_16, 7, UB_A6, UB_A3, UB_A7, -1, false, "unet c4_1"))) return e;  // [c3_3, c2_2]
    if ((e = conv(F, n, st, UC_F_16_16, 8, UB_A7, -1, UB_A8, -1, false, "unet c4_2"))) return e;
    if ((e = convT(F, n, st, 9, UB_A8, UB_A9))) return e;                                       // c4_3
    // c5_1 over [c4_3, c1_2], with the 1x1 head applied in its epilogue
    return conv(F, n, st, UC_F_24_8_HEAD, 10, UB_A9, UB_A1, UB_A10, -1, false, "unet c5_1 + head");
}

int unet_backward_frames(const UnetFrame *F, int n, hipStream_t st)
{
    using namespace xdn;
    int e = 0;
    if ((e = head_bwd(F, n, st))) return e;                                      // dY10, set 12
    if ((e = wgrad(F, n, st, UW_24_8, 11, UB_A9, UB_A1, UB_DY10))) return e;     // c5_1
    if ((e = conv(F, n, st, UC_D_8_24, 10, UB_DY10, -1, UB_DC5, -1, true, "unet c5_1 dgrad"))) return e;
    if ((e = convT_bwd(F, n, st, 9, 10, UB_A8, UB_DC5, UB_DY8))) return e;       // c4_3 <- dC5[:, 0:16]
    if ((e = wgrad(F, n, st, UW_16_16, 9, UB_A7, -1, UB_DY8))) return e;         // c4_2
    if ((e = conv(F, n, st, UC_D_16_16, 8, UB_DY8, -1, UB_DY7, UB_A7, true, "unet c4_2 dgrad"))) return e;
    if ((e = wgrad(F, n, st, UW_48_16_M0, 7, UB_A6, UB_A3, UB_DY7))) return e;   // c4_1, inputs 0-31
    if ((e = wgrad(F, n, st, UW_48_16_M1, 8, UB_A6, UB_A3, UB_DY7))) return e;   // c4_1, inputs 32-47
    if ((e = conv(F, n, st, UC_D_16_48, 7, UB_DY7, -1, UB_DC4, -1, true, "unet c4_1 dgrad"))) return e;
    if ((e = convT_bwd(F, n, st, 6, 6, UB_A5, UB_DC4, UB_DY5))) return e;        // c3_3 <- dC4[:, 0:32]
    if ((e = wgrad(F, n, st, UW_32_32, 5, UB_A4, -1, UB_DY5))) return e;         // c3_2
    if ((e = conv(F, n, st, UC_D_32_32, 5, UB_DY5, -1, UB_DY4, UB_A4, true, "unet c3_2 dgrad"))) return e;
    if ((e = wgrad(F, n, st, UW_16_32, 4, UB_P2, -1, UB_DY4))) return e;         // c3_1
    if ((e = conv(F, n, st, UC_D_32_16, 4, UB_DY4, -1, UB_DP2, -1, true, "unet c3_1 dgrad"))) return e;
    // c2_2's output feeds the second pool and the skip into c4_1 (dC4[:, 32:48])
    if ((e = pool_bwd(F, n, st, 16, UB_DP2, UB_A3, UB_DC4, 32, UB_DY3))) return e;
    if ((e = wgrad(F, n, st, UW_16_16, 3, UB_A2, -1, UB_DY3))) return e;         // c2_2
    if ((e = conv(F, n, st, UC_D_16_16, 3, UB_DY3, -1, UB_DY2, UB_A2, true, "unet c2_2 dgrad"))) return e;
    if ((e = wgrad(F, n, st, UW_8_16, 2, UB_P1, -1, UB_DY2))) return e;          // c2_1
    if ((e = conv(F, n, st, UC_D_16_8, 2, UB_DY2, -1, UB_DP1, -1, true, "unet c2_1 dgrad"))) return e;
    // c1_2's output feeds the first pool and the skip into c5_1 (dC5[:, 16:24])
    if ((e = pool_bwd(F, n, st, 8, UB_DP1, UB_A1, UB_DC5, 16, UB_DY1))) return e;
    if ((e = wgrad(F, n, st, UW_8_8, 1, UB_A0, -1, UB_DY1))) return e;           // c1_2
    if ((e = conv(F, n, st, UC_D_8_8, 1, UB_DY1, -1, UB_DY0, UB_A0, true, "unet c1_2 dgrad"))) return e;
    return wgrad(F, n, st, UW_0, 0, -1, -1, UB_DY0);                             // c1_1
}

xdn::RedJob unet_red_job(int set, float *part, float *stage, float *grad, int nblk_total)
{
    const int l = kUnetJob[set].layer;
    const UnetLayer &L = kUnet[l];
    xdn::RedJob r;
    r.part = part;
    r.stage = stage;
    r.grad = grad + kULay.w_off[l];
    r.nblk = nblk_total;
    r.ci = L.ci;
    r.co = L.co;
    r.taps = L.type == UL_CONV ? 9 : 1;
    r.transpose = L.type == UL_CONVT;  // Keras kernel [co][ci]; partials hold [ci][co]
    if (set == 7) r.ci_off = 0, r.ci_cnt = 32;               // c4_1, input channels 0-31 (+ bias)
    if (set == 8) r.ci_off = 32, r.ci_cnt = 16, r.bias = 0;  // c4_1, input channels 32-47
    r.compact = 1;
    r.pstride = unet_job_pstride(set);
    return r;
}

void unet_pack_jobs(const float *params, void *packed, std::vector<xdn::UPackJob> &jobs)
{
    char *pk = static_cast<char *>(packed);
    for (int l = 0; l < UNET_LAYERS; ++l)
    {
        const UnetLayer &L = kUnet[l];
        xdn::UPackJob j{};
        j.w = params + kULay.w_off[l];
        j.b = params + kULay.b_off[l];
        if (L.type == UL_CONV)
        {
            j.frag_fwd = pk + kULay.fwd_off[l];
            j.frag_dgrad = l > 0 ? pk + kULay.dg_off[l] : nullptr;
            j.bias32 = reinterpret_cast<float *>(pk + kULay.bias_off[l]);
            j.ci = L.ci;
            j.co = L.co;
        }
        else
        {
            j.frag_fwd = pk + kULay.master_off[l];
            j.copy_n = L.ci * L.co + L.co;  // kernel then bias, contiguous
        }
        jobs.push_back(j);
    }
}

}  // namespace xdn_abi

extern "C" {

size_t xdn_unet_param_count(void) { return kULay.count; }

size_t xdn_unet_packed_bytes(void) { return kULay.packed; }

size_t xdn_unet_workspace_size(int batch, int height, int width)
{
    if (batch <= 0 || height <= 0 || width <= 0) return 0;
    return UnetWs(nullptr, batch, height, width).bytes;
}

int xdn_unet_workspace_layout(int batch, int height, int width, int index, size_t *offset, size_t *bytes)
{
    if (batch <= 0 || height <= 0 || width <= 0 || index < 0 || index >= UNET_BUFS || !offset || !bytes)
        return fail(XDN_EINVAL, "unet workspace layout: bad arguments");
    UnetWs ws(nullptr, batch, height, width);
    *offset = ws.off[index];
    *bytes = ws.frame_bytes[index] * batch;
    return XDN_OK;
}

int xdn_unet_pack(const float *params, void *packed, void *stream)
{
    if (!params || !packed) return fail(XDN_EINVAL, "unet pack: null buffer");
    std::vector<xdn::UPackJob> jobs;
    unet_pack_jobs(params, packed, jobs);
    return hip_check(xdn::launch_upack((int)jobs.size(), jobs.data(), S(stream)), "unet pack");
}

static int unet_check(int B, int H, int W, size_t ws_bytes)
{
    if (B < 0 || H <= 0 || W <= 0) return fail(XDN_EINVAL, "unet: invalid shape [%d, %d, %d]", B, H, W);
    if (B > XDN_MAX_FRAMES) return fail(XDN_EINVAL, "unet: batch must be <= %d", XDN_MAX_FRAMES);
    // two 2x2 poolings and two 2x upsamplings must give back the full size for the concatenations
    if (H % 4 || W % 4) return fail(XDN_EINVAL, "unet: height and width must be multiples of 4 (got %d x %d)", H, W);
    if (ws_bytes < xdn_unet_workspace_size(B, H, W)) return fail(XDN_EINVAL, "unet: workspace too small");
    return XDN_OK;
}

static std::vector<UnetFrame> batch_frames(const UnetWs &ws, int B, const float *const in[5], const void *packed,
                                           float *param, const float *dparam)
{
    const size_t n0 = px(ws.h[0], ws.w[0]);
    std::vector<UnetFrame> F(B);
    for (int b = 0; b < B; ++b)
    {
        const float *img[5];
        for (int k = 0; k < 5; ++k) img[k] = in[k] + (size_t)b * n0 * 3;
        F[b] = unet_frame(ws, b, img, packed, param ? param + (size_t)b * n0 * 6 : nullptr,
                          dparam ? dparam + (size_t)b * n0 * 6 : nullptr);
    }
    return F;
}

int xdn_unet_forward(const float *ols1, const float *ols2, const float *prev_out, const float *albedo,
                     const float *normal, const void *packed, float *param, void *workspace, size_t workspace_bytes,
                     int batch, int height, int width, void *stream)
{
    if (int e = unet_check(batch, height, width, workspace_bytes)) return e;
    if (batch == 0) return XDN_OK;  // empty batch: nothing to read or write
    if (!ols1 || !ols2 || !prev_out || !albedo || !normal || !packed || !param || !workspace)
        return fail(XDN_EINVAL, "unet forward: null buffer");
    const UnetWs ws(workspace, batch, height, width);
    const float *in[5] = {ols1, ols2, prev_out, albedo, normal};
    const std::vector<UnetFrame> F = batch_frames(ws, batch, in, packed, param, nullptr);
    return unet_forward_frames(F.data(), batch, S(stream));
}

int xdn_unet_backward(const float *ols1, const float *ols2, const float *prev_out, const float *albedo,
                      const float *normal, const float *dparam, const void *packed, float *grad, void *workspace,
                      size_t workspace_bytes, int batch, int height, int width, void *stream)
{
    if (int e = unet_check(batch, height, width, workspace_bytes)) return e;
    if (batch == 0)  // empty batch: the gradient is a sum over no images
    {
        if (!grad) return fail(XDN_EINVAL, "unet backward: null buffer");
        return hip_check(xdn::launch_zero(grad, kULay.count, S(stream)), "unet backward");
    }
    if (!ols1 || !ols2 || !prev_out || !albedo || !normal || !dparam || !packed || !grad || !workspace)
        return fail(XDN_EINVAL, "unet backward: null buffer");
    const UnetWs ws(workspace, batch, height, width);
    const float *in[5] = {ols1, ols2, prev_out, albedo, normal};
    const std::vector<UnetFrame> F = batch_frames(ws, batch, in, packed, nullptr, dparam);
    if (int e = unet_backward_frames(F.data(), batch, S(stream))) return e;
    // the B images' partials of each set are contiguous: one job sums them into the one gradient
    std::vector<xdn::RedJob> jobs;
    for (int j = 0; j < UNET_JOBS; ++j) jobs.push_back(unet_red_job(j, ws.part[j], ws.stage[j], grad, batch * ws.nblk[j]));
    return hip_check(xdn::launch_wgrad_reduce((int)jobs.size(), jobs.data(), S(stream)), "unet wgrad reduce");
}

}  // extern "C"
