// Batched per-frame online denoising step over many render streams (xdn_denoise_step).
//
// One call advances every given slot (stream) by one frame. Each stage of train_step
// (scripts/main.py:250-313) is ONE launch over all slots -- the frame-batch tile maps of the
// kernels handle different resolutions in the same launch -- so n streams cost as many launches
// as one: cross-regression (3), reprojection, erosion + screening, network forward (cnn5: layer 1,
// then layers 2-5 in one row-streaming launch; the U-Net: 14 launches), fused filter + loss + backward, loss sums, network
// backward (cnn5: 5 launches, the weight and data gradients of layers 2-5 fused; the
// U-Net: 23 launches), one pair of
// weight-gradient reductions, Adam, pack.
// Every slot owns its history, network, Adam moments and scratch; nothing is shared.
#include <math.h>

#include <vector>

#include "abi_internal.h"
#include "xdn_kernels.h"
#include "xdn_misc.h"

using namespace xdn_abi;

namespace {

constexpr int kRegWin = 17, kSparsity = 4, kFilterWin = 11;  // scripts/main.py:71-73
constexpr float kAdamB1 = 0.9f, kAdamB2 = 0.999f;              // Keras Adam defaults (main.py:81)

// Scratch layout of one slot (all offsets 256-byte aligned).
struct SlotScratch {
    float *y1, *y2, *var1, *var2, *beta, *rp[3], *prev_out, *prev_ols1, *prev_ols2, *success_prev, *succ_e;
    float *dparam, *param, *ones, *loss_parts, *adam_alpha;
    uint8_t *success;
    void *cnn_ws;
    size_t cnn_ws_bytes, beta_bytes, bytes;
    SlotScratch(void *base, int H, int W)
    {
        char *p = static_cast<char *>(base);
        size_t o = 0;
        const size_t n = px(H, W);
        auto take = [&](size_t bytes) {
            char *r = p + o;
            o += al256(bytes);
            return r;
        };
        y1 = (float *)take(n * 12); y2 = (float *)take(n * 12);
        var1 = (float *)take(n * 4); var2 = (float *)take(n * 4);
        beta_bytes = xdn_cross_regression_workspace_size(1, H, W, kSparsity);
        beta = (float *)take(beta_bytes);
        success = (uint8_t *)take(n);
        for (int i = 0; i < 3; ++i) rp[i] = (float *)take(n * 12);
        prev_out = (float *)take(n * 12); prev_ols1 = (float *)take(n * 12); prev_ols2 = (float *)take(n * 12);
        success_prev = (float *)take(n * 4); succ_e = (float *)take(n * 4);
        dparam = (float *)take(n * 24); param = (float *)take(n * 24);
        ones = (float *)take(n * 4);
        loss_parts = (float *)take((size_t)xdn::filter_train_tiles(H, W) * 4);
        // either network's workspace (one slot can be admitted with either)
        const size_t c5 = Cnn5Ws(nullptr, 1, H, W).bytes, un = UnetWs(nullptr, 1, H, W).bytes;
        cnn_ws_bytes = c5 > un ? c5 : un;
        cnn_ws = take(cnn_ws_bytes);
        adam_alpha = (float *)take(4);
        bytes = o;
    }
};

__global__ void fill_kernel(float *p, size_t n, float v)
{
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}

int check_slot(const xdn_slot &s)
{
    if (s.height <= kRegWin / 2 || s.width <= kRegWin / 2)
        return fail(XDN_EINVAL, "slot size %dx%d too small", s.height, s.width);
    if (s.height % kSparsity || s.width % kSparsity)
        return fail(XDN_EINVAL, "Image height and width must be divisible by sparsity");
    if (s.network != XDN_NET_CNN5 && s.network != XDN_NET_UNET) return fail(XDN_EINVAL, "unknown network %d", s.network);
    if ((size_t)s.height * s.width >= kFilterMaxPx)
        return fail(XDN_EINVAL, "slots hold frames of fewer than 2^32 pixels (got %d x %d)", s.height, s.width);
    if (s.network == XDN_NET_CNN5 && (size_t)s.height * s.width >= kCnn5MaxPx)
        return fail(XDN_EINVAL, "cnn5 slots hold frames of fewer than 2^25 pixels (got %d x %d)", s.height, s.width);
    return XDN_OK;
}

// The buffers the requested stages read or write must be set (a null one would fault the device).
int check_slot_ptrs(const xdn_slot &s, int stages)
{
    struct Need {
        const void *p;
        int stages;
        const char *name;
    };
    constexpr int R = XDN_STAGE_REGRESS, P = XDN_STAGE_REPROJECT, F = XDN_STAGE_FINISH;
    const Need need[] = {
        {s.scratch, R | P | F, "scratch"},   {s.color, R, "color"},           {s.color2, R, "color2"},
        {s.albedo, R | F, "albedo"},         {s.normal, R | P | F, "normal"}, {s.ols1, R | F, "ols1"},
        {s.ols2, R | F, "ols2"},             {s.mvec, P, "mvec"},             {s.linear_z, P, "linear_z"},
        {s.prev_linear_z, P, "prev_linear_z"}, {s.prev_normal, P, "prev_normal"}, {s.pn_fwidth, P, "pn_fwidth"},
        {s.prev_out_log, P, "prev_out_log"}, {s.prev_ols1, P, "prev_ols1"}, {s.prev_ols2, P, "prev_ols2"},
        {s.opacity, F, "opacity"},           {s.out_log, F, "out_log"},       {s.out, F, "out"},
        {s.params, F, "params"},             {s.grad, F, "grad"},             {s.adam_m, F, "adam_m"},
        {s.adam_v, F, "adam_v"},             {s.packed, F, "packed"},         {s.avg_loss, F, "avg_loss"},
    };
    for (const Need &n : need)
        if ((n.stages & stages) && !n.p) return fail(XDN_EINVAL, "slot buffer %s is null", n.name);
    return XDN_OK;
}

}  // namespace

extern "C" {

size_t xdn_slot_scratch_bytes(int height, int width)
{
    if (height <= 0 || width <= 0) return 0;
    return SlotScratch(nullptr, height, width).bytes;
}

int xdn_slot_scratch_offsets(int height, int width, size_t *offsets)
{
    if (height <= 0 || width <= 0) return fail(XDN_EINVAL, "invalid slot size");
    char *base = reinterpret_cast<char *>((size_t)1 << 40);  // any aligned base: offsets are relative
    SlotScratch sc(base, height, width);
    const void *p[XDN_SLOT_NBUF] = {sc.y1, sc.y2, sc.var1, sc.var2, sc.beta, sc.success, sc.rp[0], sc.rp[1],
                                    sc.rp[2], sc.prev_out, sc.prev_ols1, sc.prev_ols2, sc.success_prev, sc.succ_e,
                                    sc.dparam, sc.param, sc.ones, sc.loss_parts, sc.cnn_ws, sc.adam_alpha};
    for (int i = 0; i < XDN_SLOT_NBUF; ++i) offsets[i] = (size_t)(static_cast<const char *>(p[i]) - base);
    return XDN_OK;
}

int xdn_slot_init(const xdn_slot *slot, void *stream)
{
    if (int e = check_slot(*slot)) return e;
    if (!slot->scratch || !slot->params || !slot->packed)
        return fail(XDN_EINVAL, "slot init needs scratch, params and packed");
    SlotScratch sc(slot->scratch, slot->height, slot->width);
    const size_t n = px(slot->height, slot->width);
    hipLaunchKernelGGL(fill_kernel, dim3(256), dim3(256), 0, S(stream), sc.ones, n, 1.0f);
    if (int e = hip_check(hipGetLastError(), "slot init")) return e;
    return slot->network == XDN_NET_UNET ? xdn_unet_pack(slot->params, slot->packed, stream)
                                         : xdn_cnn5_pack(slot->params, slot->packed, stream);
}

int xdn_denoise_step(const xdn_slot *slots, int n_slots, float lr, void *stream)
{
    return xdn_denoise_stages(slots, n_slots, lr, XDN_STAGE_ALL, stream);
}

int xdn_denoise_stages(const xdn_slot *slots, int n_slots, float lr, int stages, void *stream)
{
    if (n_slots < 0 || n_slots > XDN_MAX_FRAMES) return fail(XDN_EINVAL, "n_slots must be in [0, %d]", XDN_MAX_FRAMES);
    if (stages & ~XDN_STAGE_ALL) return fail(XDN_EINVAL, "unknown stage bits 0x%x", stages);
    for (int i = 0; i < n_slots; ++i)
    {
        if (int e = check_slot(slots[i])) return e;
        if (slots[i].network != slots[0].network) return fail(XDN_EINVAL, "one call takes slots of one network");
    }
    for (int i = 0; i < n_slots; ++i)
        if (int e = check_slot_ptrs(slots[i], stages)) return e;
    if (n_slots == 0) return XDN_OK;
    const bool unet = slots[0].network == XDN_NET_UNET;
    hipStream_t st = S(stream);
    const int n = n_slots;
    std::vector<SlotScratch> sc;
    std::vector<Cnn5Ws> cw;
    std::vector<UnetWs> uw;
    std::vector<UnetFrame> uf;
    sc.reserve(n);
    (unet ? uw.reserve(n) : cw.reserve(n));
    for (int i = 0; i < n; ++i)
    {
        const xdn_slot &s = slots[i];
        sc.emplace_back(s.scratch, s.height, s.width);
        if (unet)
        {
            uw.emplace_back(sc[i].cnn_ws, 1, s.height, s.width);
            const float *img[5] = {s.ols1, s.ols2, sc[i].prev_out, s.albedo, s.normal};
            uf.push_back(unet_frame(uw[i], 0, img, s.packed, sc[i].param, sc[i].dparam));
        }
        else
            cw.emplace_back(sc[i].cnn_ws, 1, s.height, s.width);
    }
    // 1. cross-regression with the log1p prologue (variance launches hold 2 frames per slot)
    if (stages & XDN_STAGE_REGRESS)
    {
        std::vector<xdn::CrFrame> fr(n);
        for (int i = 0; i < n; ++i)
        {
            const xdn_slot &s = slots[i];
            xdn::CrFrame &f = fr[i];
            f.color = s.color; f.color2 = s.color2; f.y1 = sc[i].y1; f.y2 = sc[i].y2;
            f.var1 = sc[i].var1; f.var2 = sc[i].var2; f.alb = s.albedo; f.nrm = s.normal; f.beta = sc[i].beta;
            f.outa = s.ols1; f.outb = s.ols2; f.H = s.height; f.W = s.width;
        }
        for (int b0 = 0; b0 < n; b0 += XDN_MAX_FRAMES / 2)
        {
            const int m = n - b0 < XDN_MAX_FRAMES / 2 ? n - b0 : XDN_MAX_FRAMES / 2;
            if (int e = hip_check(xdn::launch_variance(m, fr.data() + b0, 3, true, st), "variance")) return e;
        }
        if (int e = hip_check(xdn::launch_regression(n, fr.data(), kRegWin, kSparsity, st), "regression")) return e;
        if (int e = hip_check(xdn::launch_recon(n, fr.data(), kRegWin, kSparsity, 1, st), "recon")) return e;
    }
    // 2. reprojection of [out_log, ols1, ols2] of the previous frame (main.py:266, opacities one)
    if (stages & XDN_STAGE_REPROJECT)
    {
        std::vector<xdn::RpFrame> fr(n);
        const int dims[3] = {3, 3, 3};
        for (int i = 0; i < n; ++i)
        {
            const xdn_slot &s = slots[i];
            xdn::RpFrame &f = fr[i];
            memset(&f, 0, sizeof(f));
            f.mvec = s.mvec; f.linZ = s.linear_z; f.prevLinZ = s.prev_linear_z; f.nrm = s.normal;
            f.prevNrm = s.prev_normal; f.pnFw = s.pn_fwidth; f.opac = sc[i].ones; f.prevOpac = sc[i].ones;
            f.success = sc[i].success;
            f.in[0] = s.prev_out_log; f.in[1] = s.prev_ols1; f.in[2] = s.prev_ols2;
            f.out[0] = sc[i].rp[0]; f.out[1] = sc[i].rp[1]; f.out[2] = sc[i].rp[2];
            f.H = s.height; f.W = s.width;
        }
        if (int e = hip_check(xdn::launch_reproject(n, fr.data(), 3, dims, true, st), "reproject")) return e;
    }
    if (!(stages & XDN_STAGE_FINISH)) return XDN_OK;
    // 3. 3x3 erosion + temporal screening (win_curr 7, win_prev 1, main.py:270-280)
    {
        std::vector<xdn::TsFrame> fr(n);
        for (int i = 0; i < n; ++i)
        {
            const xdn_slot &s = slots[i];
            xdn::TsFrame &f = fr[i];
            f.c1 = s.ols1; f.c2 = s.ols2; f.prev = sc[i].rp[0]; f.prev1 = sc[i].rp[1]; f.prev2 = sc[i].rp[2];
            f.rsucc = nullptr; f.succ_mask = sc[i].success; f.succ_e = sc[i].succ_e;
            f.o = sc[i].prev_out; f.o1 = sc[i].prev_ols1; f.o2 = sc[i].prev_ols2; f.osucc = sc[i].success_prev;
            f.H = s.height; f.W = s.width;
        }
        if (int e = hip_check(xdn::launch_screening(n, fr.data(), 7, 1, true, st), "screening")) return e;
    }
    // 4. network forward on [ols1, ols2, prev_out, albedo, normal] (main.py:291)
    if (unet)
    {
        if (int e = unet_forward_frames(uf.data(), n, st)) return e;
    }
    else
    {
        // layer 1 (the five FP32 images in), then layers 2..5 in one row-streaming launch
        std::vector<xdn::ConvFrame> fr(n);
        std::vector<xdn::FwdStreamFrame> sf(n);
        for (int i = 0; i < n; ++i)
        {
            const xdn_slot &s = slots[i];
            const char *pk = static_cast<const char *>(s.packed);
            xdn::ConvFrame &f = fr[i];
            memset(&f, 0, sizeof(f));
            f.in[0] = s.ols1; f.in[1] = s.ols2; f.in[2] = sc[i].prev_out; f.in[3] = s.albedo; f.in[4] = s.normal;
            f.out = cw[i].X[0];
            f.wfrag = pk + kLay.fwd_off[0];
            f.bias = reinterpret_cast<const float *>(pk + kLay.bias_off[0]);
            f.H = s.height; f.W = s.width;
            sf[i] = cnn5_stream_frame(cw[i], pk, 0, sc[i].param, s.height, s.width);
        }
        if (int e = hip_check(xdn::launch_conv_l1(n, fr.data(), st), "cnn forward")) return e;
        if (int e = hip_check(xdn::launch_fwd_stream(n, sf.data(), st), "cnn forward")) return e;
    }
    // 5. fused spatial filter + temporal accumulation + loss + backward to the network output
    {
        std::vector<xdn::TrainFrame> fr(n);
        std::vector<const float *> parts(n);
        std::vector<int> cnt(n);
        std::vector<float> scale(n);
        std::vector<float *> outp(n);
        for (int i = 0; i < n; ++i)
        {
            const xdn_slot &s = slots[i];
            xdn::TrainFrame &f = fr[i];
            f.ols1 = s.ols1; f.ols2 = s.ols2; f.alb = s.albedo; f.nrm = s.normal; f.param = sc[i].param;
            f.succ_e = sc[i].succ_e; f.succ_prev = sc[i].success_prev; f.prev_out = sc[i].prev_out;
            f.opacity = s.opacity; f.prev_ols1 = sc[i].prev_ols1; f.prev_ols2 = sc[i].prev_ols2;
            f.out_log = s.out_log; f.out = s.out; f.dparam = sc[i].dparam; f.loss = nullptr;
            f.loss_part = sc[i].loss_parts; f.H = s.height; f.W = s.width;
            parts[i] = sc[i].loss_parts;
            cnt[i] = xdn::filter_train_tiles(s.height, s.width);
            scale[i] = (float)(1.0 / (3.0 * (double)px(s.height, s.width)));
            outp[i] = s.avg_loss;
        }
        // the loss-sum launch also advances the device Adam step counters used by step 7
        xdn::AdamTick tick{};
        tick.lr = lr; tick.b1 = kAdamB1; tick.b2 = kAdamB2;
        for (int i = 0; i < n; ++i)
            if (slots[i].adam_t)
            {
                tick.counter[tick.n] = slots[i].adam_t;
                tick.alpha[tick.n] = sc[i].adam_alpha;
                ++tick.n;
            }
        if (int e = hip_check(xdn::launch_filter_train(n, fr.data(), kFilterWin, st), "filter_train")) return e;
        if (int e = hip_check(xdn::launch_sums(n, parts.data(), cnt.data(), scale.data(), outp.data(), st, &tick),
                              "loss"))
            return e;
    }
    // 6. network backward: layers 5..3 each in one fused launch (weight gradient and data gradient
    //    into the layer below), then layer 2's backward fused with layer 1's weight gradient (layer
    //    1's pre-activation gradient stays on chip); the per-block weight-gradient partials of every
    //    (slot, layer) are reduced by one launch pair
    if (unet)
    {
        if (int e = unet_backward_frames(uf.data(), n, st)) return e;
        std::vector<xdn::RedJob> jobs;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < UNET_JOBS; ++j)
                jobs.push_back(unet_red_job(j, uw[i].part[j], uw[i].stage[j], slots[i].grad, uw[i].nblk[j]));
        if (int e = hip_check(xdn::launch_wgrad_reduce((int)jobs.size(), jobs.data(), st), "unet wgrad reduce")) return e;
    }
    else
    {
        std::vector<xdn::RedJob> jobs;
        std::vector<xdn::BwdFrame> bf(n);
        int cur = 0;  // dY ping-pong index holding the gradient of the current layer's output
        for (int l = 4; l >= 2; --l)
        {
            // the last layer's dY is the FP32 gradient of the network output
            for (int i = 0; i < n; ++i)
            {
                const xdn_slot &s = slots[i];
                const void *dy = l == 4 ? static_cast<const void *>(sc[i].dparam) : static_cast<const void *>(cw[i].dY[cur]);
                bf[i] = xdn::BwdFrame{dy, cw[i].X[l - 1], cw[i].dY[1 - cur],
                                      static_cast<const char *>(s.packed) + kLay.dg_off[l], cw[i].part[l],
                                      s.height, s.width, cw[i].nblk[l]};
                jobs.push_back(cnn5_red_job(cw[i], l, s.grad, cw[i].nblk[l]));
            }
            if (int e = hip_check(xdn::launch_bwd(n, bf.data(), l == 4, st), "cnn backward")) return e;
            cur = 1 - cur;
        }
        std::vector<xdn::Bwd12Frame> f12(n);
        for (int i = 0; i < n; ++i)
        {
            const xdn_slot &s = slots[i];
            f12[i] = xdn::Bwd12Frame{cw[i].dY[cur], cw[i].X[0], static_cast<const char *>(s.packed) + kLay.dg_off[1],
                                     cw[i].part[1], {s.ols1, s.ols2, sc[i].prev_out, s.albedo, s.normal},
                                     cw[i].part[0], s.height, s.width, cw[i].nblk[1]};
            jobs.push_back(cnn5_red_job(cw[i], 1, s.grad, cw[i].nblk[1]));
            jobs.push_back(cnn5_red_job(cw[i], 0, s.grad, cw[i].nblk[0]));
        }
        if (int e = hip_check(xdn::launch_bwd12(n, f12.data(), st), "cnn backward")) return e;
        if (int e = hip_check(xdn::launch_wgrad_reduce((int)jobs.size(), jobs.data(), st), "cnn wgrad reduce")) return e;
    }
    // 7. Keras Adam on every slot's parameters and the bf16 fragments: the cnn5 update writes each
    //    weight's fragment slots itself; the U-Net's fragments are rebuilt by one packing launch
    {
        std::vector<xdn::AdamJob> jobs(n);
        std::vector<xdn::UPackJob> upj;
        xdn::AdamPackLayer lay[xdn::ADAM_PACK_LAYERS];
        for (int l = 0; l < 5; ++l)
            lay[l] = xdn::AdamPackLayer{(int)kLay.w_off[l], (int)kLay.b_off[l], kCnn5[l].ci, kCnn5[l].co,
                                        (int)kLay.fwd_off[l], l > 0 ? (int)kLay.dg_off[l] : -1,
                                        (int)kLay.bias_off[l]};
        const float b1 = kAdamB1, b2 = kAdamB2, eps = 1e-7f;
        for (int i = 0; i < n; ++i)
        {
            const xdn_slot &s = slots[i];
            const int t = s.adam_step < 1 ? 1 : s.adam_step;
            jobs[i].param = s.params; jobs[i].grad = s.grad; jobs[i].m = s.adam_m; jobs[i].v = s.adam_v;
            jobs[i].count = (int)(unet ? kULay.count : kLay.count);
            jobs[i].alpha = (float)((double)lr * sqrt(1.0 - pow((double)b2, t)) / (1.0 - pow((double)b1, t)));
            jobs[i].alpha_dev = s.adam_t ? sc[i].adam_alpha : nullptr;
            jobs[i].packed = unet ? nullptr : s.packed;
            if (unet) unet_pack_jobs(s.params, s.packed, upj);
        }
        if (int e = hip_check(xdn::launch_adam(n, jobs.data(), b1, b2, eps, st, unet ? nullptr : lay), "adam"))
            return e;
        if (unet)
            if (int e = hip_check(xdn::launch_upack((int)upj.size(), upj.data(), st), "unet pack")) return e;
    }
    return XDN_OK;
}

int xdn_denoise_graph_create(const xdn_slot *slots, int n_slots, float lr, void *stream, void **graph)
{
    if (!graph) return fail(XDN_EINVAL, "graph output pointer is null");
    *graph = nullptr;
    if (!stream) return fail(XDN_EINVAL, "graph capture needs a non-null stream");
    for (int i = 0; i < n_slots; ++i)
        if (!slots[i].adam_t) return fail(XDN_EINVAL, "graph capture needs adam_t (device step counter) on every slot");
    hipStream_t st = S(stream);
    if (int e = hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture")) return e;
    const int rc = xdn_denoise_step(slots, n_slots, lr, stream);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &g);
    if (rc != XDN_OK)
    {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (int e = hip_check(ec, "end capture")) return e;
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (int e = hip_check(ei, "graph instantiate")) return e;
    *graph = ex;
    return XDN_OK;
}

int xdn_denoise_graph_launch(void *graph, void *stream)
{
    if (!graph) return fail(XDN_EINVAL, "null graph");
    return hip_check(hipGraphLaunch(static_cast<hipGraphExec_t>(graph), S(stream)), "graph launch");
}

int xdn_denoise_graph_destroy(void *graph)
{
    if (!graph) return XDN_OK;
    return hip_check(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph)), "graph destroy");
}

}  // extern "C"
