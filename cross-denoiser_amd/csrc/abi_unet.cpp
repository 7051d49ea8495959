// extern "C" entry points of the U-Net (declared in include/xdn.h): the reference's bandwidth
// network, scripts/main.py:83-118, forward and backward as launches of the k_unet.hip kernels.
#include <hip/hip_runtime.h>
#include <string.h>

#include <vector>

#include "abi_internal.h"
#include "xdn_unet.h"

using namespace xdn_abi;

namespace {

struct Ctx {
    const UnetWs &ws;
    const char *pk;
    const float *img[5];
    int B;
    hipStream_t st;
    size_t n0() const { return (size_t)ws.h[0] * ws.w[0]; }
};

const void *pk_at(const Ctx &c, size_t off) { return c.pk + off; }

// 3x3 convolution of layer l: forward (dgrad = false) from concat[a, b] (or the images when a < 0)
// into buffer out; data gradient from buffer a into out, masked by buffer mask (or FP32 when
// mask < 0).
int conv(const Ctx &c, int kind, int l, int a, int b, int out, int mask, bool dgrad, const char *what)
{
    const int lv = dgrad ? kUnetBuf[out].level : kUnet[l].level;
    std::vector<xdn::UcFrame> fr(c.B);
    for (int i = 0; i < c.B; ++i)
    {
        xdn::UcFrame &f = fr[i];
        memset(&f, 0, sizeof(f));
        if (a >= 0)
            f.a = c.ws.at(a, i);
        else
            for (int k = 0; k < 5; ++k) f.img[k] = c.img[k] + (size_t)i * c.n0() * 3;
        f.b = b >= 0 ? c.ws.at(b, i) : nullptr;
        f.out = c.ws.at(out, i);
        f.mask = mask >= 0 ? static_cast<const uint16_t *>(c.ws.at(mask, i)) : nullptr;
        f.wfrag = pk_at(c, dgrad ? kULay.dg_off[l] : kULay.fwd_off[l]);
        f.bias = reinterpret_cast<const float *>(pk_at(c, kULay.bias_off[l]));
        f.H = c.ws.h[lv];
        f.W = c.ws.w[lv];
    }
    return hip_check(xdn::launch_uconv(kind, c.B, fr.data(), c.st), what);
}

int pool(const Ctx &c, int C, int in, int out)
{
    const int lv = kUnetBuf[in].level;
    std::vector<xdn::PoolFrame> fr(c.B);
    for (int i = 0; i < c.B; ++i)
        fr[i] = xdn::PoolFrame{static_cast<const uint16_t *>(c.ws.at(in, i)), static_cast<uint16_t *>(c.ws.at(out, i)),
                               c.ws.h[lv], c.ws.w[lv]};
    return hip_check(xdn::launch_maxpool(C, c.B, fr.data(), c.st), "unet max pool");
}

const float *master(const Ctx &c, int l) { return reinterpret_cast<const float *>(pk_at(c, kULay.master_off[l])); }

}  // namespace

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
    char *pk = static_cast<char *>(packed);
    std::vector<xdn::UPackJob> jobs;
    for (int l = 0; l < UNET_LAYERS; ++l)
    {
        const UnetLayer &L = kUnet[l];
        if (L.type == UL_CONV)
        {
            xdn::UPackJob j{};
            j.w = params + kULay.w_off[l];
            j.b = params + kULay.b_off[l];
            j.frag_fwd = pk + kULay.fwd_off[l];
            j.frag_dgrad = l > 0 ? pk + kULay.dg_off[l] : nullptr;
            j.bias32 = reinterpret_cast<float *>(pk + kULay.bias_off[l]);
            j.ci = L.ci;
            j.co = L.co;
            jobs.push_back(j);
        }
        else
        {
            const size_t n = (size_t)L.ci * L.co + L.co;  // kernel then bias, contiguous
            if (int e = hip_check(hipMemcpyAsync(pk + kULay.master_off[l], params + kULay.w_off[l], n * 4,
                                                 hipMemcpyDeviceToDevice, S(stream)),
                                  "unet pack copy"))
                return e;
        }
    }
    return hip_check(xdn::launch_upack((int)jobs.size(), jobs.data(), S(stream)), "unet pack");
}

static int unet_check(int B, int H, int W, size_t ws_bytes)
{
    if (B <= 0 || H <= 0 || W <= 0) return fail(XDN_EINVAL, "unet: invalid shape [%d, %d, %d]", B, H, W);
    if (B > XDN_MAX_FRAMES) return fail(XDN_EINVAL, "unet: batch must be <= %d", XDN_MAX_FRAMES);
    // two 2x2 poolings and two 2x upsamplings must give back the full size for the concatenations
    if (H % 4 || W % 4) return fail(XDN_EINVAL, "unet: height and width must be multiples of 4 (got %d x %d)", H, W);
    if (ws_bytes < xdn_unet_workspace_size(B, H, W)) return fail(XDN_EINVAL, "unet: workspace too small");
    return XDN_OK;
}

int xdn_unet_forward(const float *ols1, const float *ols2, const float *prev_out, const float *albedo,
                     const float *normal, const void *packed, float *param, void *workspace, size_t workspace_bytes,
                     int batch, int height, int width, void *stream)
{
    if (int e = unet_check(batch, height, width, workspace_bytes)) return e;
    const UnetWs ws(workspace, batch, height, width);
    const Ctx c{ws, static_cast<const char *>(packed), {ols1, ols2, prev_out, albedo, normal}, batch, S(stream)};
    using namespace xdn;
    int e = 0;
    if ((e = conv(c, UC_F_0, 0, -1, -1, UB_A0, -1, false, "unet c1_1"))) return e;
    if ((e = conv(c, UC_F_8_8, 1, UB_A0, -1, UB_A1, -1, false, "unet c1_2"))) return e;
    if ((e = pool(c, 8, UB_A1, UB_P1))) return e;
    if ((e = conv(c, UC_F_8_16, 2, UB_P1, -1, UB_A2, -1, false, "unet c2_1"))) return e;
    if ((e = conv(c, UC_F_16_16, 3, UB_A2, -1, UB_A3, -1, false, "unet c2_2"))) return e;
    if ((e = pool(c, 16, UB_A3, UB_P2))) return e;
    if ((e = conv(c, UC_F_16_32, 4, UB_P2, -1, UB_A4, -1, false, "unet c3_1"))) return e;
    if ((e = conv(c, UC_F_32_32, 5, UB_A4, -1, UB_A5, -1, false, "unet c3_2"))) return e;
    std::vector<CtFrame> ct(batch);
    for (int l : {6, 9})
    {
        // c3_3: A5 (quarter) -> A6 (half); c4_3: A8 (half) -> A9 (full)
        const int in = l == 6 ? UB_A5 : UB_A8, out = l == 6 ? UB_A6 : UB_A9, lv = kUnet[l].level;
        for (int i = 0; i < batch; ++i)
        {
            CtFrame &f = ct[i];
            memset(&f, 0, sizeof(f));
            f.in = static_cast<const uint16_t *>(ws.at(in, i));
            f.out = static_cast<uint16_t *>(ws.at(out, i));
            f.w = master(c, l);
            f.b = f.w + kUnet[l].ci * kUnet[l].co;
            f.h = ws.h[lv];
            f.w_ = ws.w[lv];
        }
        if ((e = hip_check(launch_convT(kUnet[l].co, batch, ct.data(), c.st), "unet transposed conv"))) return e;
        if (l == 6)
        {
            if ((e = conv(c, UC_F_48_16, 7, UB_A6, UB_A3, UB_A7, -1, false, "unet c4_1"))) return e;
            if ((e = conv(c, UC_F_16_16, 8, UB_A7, -1, UB_A8, -1, false, "unet c4_2"))) return e;
        }
    }
    if ((e = conv(c, UC_F_24_8, 10, UB_A9, UB_A1, UB_A10, -1, false, "unet c5_1"))) return e;
    std::vector<HeadFrame> hf(batch);
    for (int i = 0; i < batch; ++i)
    {
        HeadFrame &f = hf[i];
        memset(&f, 0, sizeof(f));
        f.in = static_cast<const uint16_t *>(ws.at(UB_A10, i));
        f.param = param + (size_t)i * c.n0() * 6;
        f.w = master(c, 11);
        f.b = f.w + 48;
        f.H = height;
        f.W = width;
    }
    return hip_check(launch_head(batch, hf.data(), c.st), "unet head");
}

int xdn_unet_backward(const float *ols1, const float *ols2, const float *prev_out, const float *albedo,
                      const float *normal, const float *dparam, const void *packed, float *grad, void *workspace,
                      size_t workspace_bytes, int batch, int height, int width, void *stream)
{
    if (int e = unet_check(batch, height, width, workspace_bytes)) return e;
    const UnetWs ws(workspace, batch, height, width);
    const Ctx c{ws, static_cast<const char *>(packed), {ols1, ols2, prev_out, albedo, normal}, batch, S(stream)};
    using namespace xdn;
    int e = 0;
    std::vector<RedJob> jobs;
    auto red = [&](int j, int taps, int ci_off, int ci_cnt, int bias, int transpose) {
        const int l = kUnetJob[j].layer;
        RedJob r;
        r.part = ws.part[j];
        r.stage = ws.stage[j];
        r.grad = grad + kULay.w_off[l];
        r.nblk = batch * ws.nblk[j];
        r.ci = kUnet[l].ci;
        r.co = kUnet[l].co;
        r.taps = taps;
        r.ci_off = ci_off;
        r.ci_cnt = ci_cnt;
        r.bias = bias;
        r.transpose = transpose;
        jobs.push_back(r);
    };
    // weight gradient of partial set j over concat[a, b] (or the images) and dY buffer dy
    auto wgrad = [&](int kind, int j, int a, int b, int dy) {
        const int lv = kUnetJob[j].level;
        std::vector<UwFrame> fr(batch);
        for (int i = 0; i < batch; ++i)
        {
            UwFrame &f = fr[i];
            memset(&f, 0, sizeof(f));
            if (a >= 0)
                f.a = ws.at(a, i);
            else
                for (int k = 0; k < 5; ++k) f.img[k] = c.img[k] + (size_t)i * c.n0() * 3;
            f.b = b >= 0 ? ws.at(b, i) : nullptr;
            f.dy = static_cast<const uint16_t *>(ws.at(dy, i));
            f.part = ws.part[j] + (size_t)i * ws.nblk[j] * WG_PART;
            f.nblk = ws.nblk[j];
            f.H = ws.h[lv];
            f.W = ws.w[lv];
        }
        return hip_check(launch_uwgrad(kind, batch, fr.data(), c.st), "unet wgrad");
    };
    // transposed layer l (partial set j): input buffer in, gradient of its output in FP32 buffer
    // dout (first C of its channels), gradient into its input -> din (masked by in's ReLU)
    auto convT_bwd = [&](int l, int j, int in, int dout, int din) {
        const int lv = kUnet[l].level;
        std::vector<CtFrame> fr(batch);
        for (int i = 0; i < batch; ++i)
        {
            CtFrame &f = fr[i];
            memset(&f, 0, sizeof(f));
            f.in = static_cast<const uint16_t *>(ws.at(in, i));
            f.dout = static_cast<const float *>(ws.at(dout, i));
            f.din = static_cast<uint16_t *>(ws.at(din, i));
            f.w = master(c, l);
            f.b = f.w + kUnet[l].ci * kUnet[l].co;
            f.part = ws.part[j] + (size_t)i * ws.nblk[j] * WG_PART;
            f.h = ws.h[lv];
            f.w_ = ws.w[lv];
            f.dstride = kUnetBuf[dout].ch;
            f.nblk = ws.nblk[j];
        }
        return hip_check(launch_convT_bwd(kUnet[l].co, batch, fr.data(), c.st), "unet transposed conv backward");
    };
    // gradient into a pooled activation: pooled gradient dp + skip gradient (channels soff.. of
    // buffer skip) -> dy, masked by act's ReLU
    auto pool_bwd = [&](int C, int dp, int act, int skip, int soff, int dy) {
        const int lv = kUnetBuf[act].level;
        std::vector<PoolBwFrame> fr(batch);
        for (int i = 0; i < batch; ++i)
            fr[i] = PoolBwFrame{static_cast<const float *>(ws.at(dp, i)), static_cast<const uint16_t *>(ws.at(act, i)),
                                static_cast<const float *>(ws.at(skip, i)), static_cast<uint16_t *>(ws.at(dy, i)),
                                ws.h[lv], ws.w[lv], kUnetBuf[skip].ch, soff};
        return hip_check(launch_pool_bwd(C, batch, fr.data(), c.st), "unet max pool backward");
    };

    // head: dY10 = ReLU'(A10) * W dparam; partial set 12
    {
        std::vector<HeadFrame> hf(batch);
        for (int i = 0; i < batch; ++i)
        {
            HeadFrame &f = hf[i];
            memset(&f, 0, sizeof(f));
            f.in = static_cast<const uint16_t *>(ws.at(UB_A10, i));
            f.dparam = dparam + (size_t)i * c.n0() * 6;
            f.din = static_cast<uint16_t *>(ws.at(UB_DY10, i));
            f.w = master(c, 11);
            f.b = f.w + 48;
            f.part = ws.part[12] + (size_t)i * ws.nblk[12] * WG_PART;
            f.H = height;
            f.W = width;
            f.nblk = ws.nblk[12];
        }
        if ((e = hip_check(launch_head_bwd(batch, hf.data(), c.st), "unet head backward"))) return e;
        red(12, 1, 0, -1, 1, 0);
    }
    // c5_1 over concat[A9, A1]
    if ((e = wgrad(UW_24_8, 11, UB_A9, UB_A1, UB_DY10))) return e;
    red(11, 9, 0, -1, 1, 0);
    if ((e = conv(c, UC_D_8_24, 10, UB_DY10, -1, UB_DC5, -1, true, "unet c5_1 dgrad"))) return e;
    // c4_3: gradient of its output = dC5[:, 0:16]
    if ((e = convT_bwd(9, 10, UB_A8, UB_DC5, UB_DY8))) return e;
    red(10, 1, 0, -1, 1, 1);
    // c4_2
    if ((e = wgrad(UW_16_16, 9, UB_A7, -1, UB_DY8))) return e;
    red(9, 9, 0, -1, 1, 0);
    if ((e = conv(c, UC_D_16_16, 8, UB_DY8, -1, UB_DY7, UB_A7, true, "unet c4_2 dgrad"))) return e;
    // c4_1 over concat[A6, A3]: two 32-channel input slices
    if ((e = wgrad(UW_48_16_M0, 7, UB_A6, UB_A3, UB_DY7))) return e;
    red(7, 9, 0, 32, 1, 0);
    if ((e = wgrad(UW_48_16_M1, 8, UB_A6, UB_A3, UB_DY7))) return e;
    red(8, 9, 32, 16, 0, 0);
    if ((e = conv(c, UC_D_16_48, 7, UB_DY7, -1, UB_DC4, -1, true, "unet c4_1 dgrad"))) return e;
    // c3_3: gradient of its output = dC4[:, 0:32]
    if ((e = convT_bwd(6, 6, UB_A5, UB_DC4, UB_DY5))) return e;
    red(6, 1, 0, -1, 1, 1);
    // c3_2, c3_1
    if ((e = wgrad(UW_32_32, 5, UB_A4, -1, UB_DY5))) return e;
    red(5, 9, 0, -1, 1, 0);
    if ((e = conv(c, UC_D_32_32, 5, UB_DY5, -1, UB_DY4, UB_A4, true, "unet c3_2 dgrad"))) return e;
    if ((e = wgrad(UW_16_32, 4, UB_P2, -1, UB_DY4))) return e;
    red(4, 9, 0, -1, 1, 0);
    if ((e = conv(c, UC_D_32_16, 4, UB_DY4, -1, UB_DP2, -1, true, "unet c3_1 dgrad"))) return e;
    // second pool: A3 receives the pooled gradient and the skip gradient dC4[:, 32:48]
    if ((e = pool_bwd(16, UB_DP2, UB_A3, UB_DC4, 32, UB_DY3))) return e;
    // c2_2, c2_1
    if ((e = wgrad(UW_16_16, 3, UB_A2, -1, UB_DY3))) return e;
    red(3, 9, 0, -1, 1, 0);
    if ((e = conv(c, UC_D_16_16, 3, UB_DY3, -1, UB_DY2, UB_A2, true, "unet c2_2 dgrad"))) return e;
    if ((e = wgrad(UW_8_16, 2, UB_P1, -1, UB_DY2))) return e;
    red(2, 9, 0, -1, 1, 0);
    if ((e = conv(c, UC_D_16_8, 2, UB_DY2, -1, UB_DP1, -1, true, "unet c2_1 dgrad"))) return e;
    // first pool: A1 receives the pooled gradient and the skip gradient dC5[:, 16:24]
    if ((e = pool_bwd(8, UB_DP1, UB_A1, UB_DC5, 16, UB_DY1))) return e;
    // c1_2, c1_1
    if ((e = wgrad(UW_8_8, 1, UB_A0, -1, UB_DY1))) return e;
    red(1, 9, 0, -1, 1, 0);
    if ((e = conv(c, UC_D_8_8, 1, UB_DY1, -1, UB_DY0, UB_A0, true, "unet c1_2 dgrad"))) return e;
    if ((e = wgrad(UW_0, 0, -1, -1, UB_DY0))) return e;
    red(0, 9, 0, -1, 1, 0);
    return hip_check(launch_wgrad_reduce((int)jobs.size(), jobs.data(), c.st), "unet wgrad reduce");
}

}  // extern "C"
