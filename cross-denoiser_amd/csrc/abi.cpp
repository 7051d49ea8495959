// extern "C" entry points of libxdn.so (declared in include/xdn.h): argument validation, batch
// splitting into launches of at most XDN_MAX_FRAMES frames, error reporting.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "abi_internal.h"
#include "xdn_kernels.h"
#include "xdn_misc.h"

using namespace xdn_abi;

namespace xdn_abi {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace xdn_abi

namespace {

template <class Fill, class Launch>
int sf_run(int B, int H, int W, int win, Fill fill, Launch launch, const char *what)
{
    for (int b0 = 0; b0 < B; b0 += XDN_MAX_FRAMES)
    {
        const int n = B - b0 < XDN_MAX_FRAMES ? B - b0 : XDN_MAX_FRAMES;
        std::vector<xdn::SfFrame> fr(n);
        for (int i = 0; i < n; ++i)
        {
            memset(&fr[i], 0, sizeof(xdn::SfFrame));
            fill(fr[i], (size_t)(b0 + i));
            fr[i].H = H;
            fr[i].W = W;
        }
        if (int e = hip_check(launch(n, fr.data()), what)) return e;
    }
    return XDN_OK;
}

}  // namespace

extern "C" {

const char *xdn_version(void) { return "xdn 0.1 gfx950"; }

const char *xdn_last_error(void) { return xdn_abi::g_err; }

size_t xdn_cross_regression_workspace_size(int batch, int height, int width, int sparsity)
{
    if (batch <= 0 || height <= 0 || width <= 0 || sparsity <= 0) return 0;
    return (size_t)batch * (size_t)(height / sparsity) * (size_t)(width / sparsity) * 36 * sizeof(float);
}

static int cross_regression_impl(const float *color, const float *color2, const float *imga, const float *imgb,
                                 const float *albedo, const float *normal, float *y1, float *y2, float *outa,
                                 float *outb, float *vara, float *varb, void *ws, size_t ws_bytes, int B, int H,
                                 int W, int win, int s, int rs, void *stream)
{
    if (int e = check_shape(B, H, W)) return e;
    if (win < 1 || s < 1 || rs < 1) return fail(XDN_EINVAL, "win_size, sparsity and recon_scale must be >= 1");
    if (rs > s) return fail(XDN_EINVAL, "recon_scale must be less than or equal to sparsity");
    if (H % rs != 0 || W % rs != 0) return fail(XDN_EINVAL, "Image height and width must be divisible by recon_scale");
    if (H % s != 0 || W % s != 0) return fail(XDN_EINVAL, "Image height and width must be divisible by sparsity");
    if (H <= win / 2 || W <= win / 2) return fail(XDN_EINVAL, "Image height and width must exceed win_size/2");
    if (B == 0) return XDN_OK;  // empty batch: nothing to read or write
    if ((!color && (!imga || !imgb)) || !albedo || !normal || !outa || !outb || !vara || !varb)
        return fail(XDN_EINVAL, "cross_regression: null tensor");
    if (!ws || ws_bytes < xdn_cross_regression_workspace_size(B, H, W, s))
        return fail(XDN_EINVAL, "workspace too small: %zu < %zu", ws_bytes, xdn_cross_regression_workspace_size(B, H, W, s));
    const bool fused = color != nullptr;
    const size_t N = px(H, W), Nr = px(H / rs, W / rs), Nb = (size_t)(H / s) * (W / s) * 36;
    for (int b0 = 0; b0 < B; b0 += XDN_MAX_FRAMES / 2)
    {
        const int n = B - b0 < XDN_MAX_FRAMES / 2 ? B - b0 : XDN_MAX_FRAMES / 2;
        std::vector<xdn::CrFrame> fr(n);
        for (int i = 0; i < n; ++i)
        {
            const size_t b = (size_t)(b0 + i);
            xdn::CrFrame &f = fr[i];
            f.color = fused ? color + b * N * 3 : nullptr;
            f.color2 = fused ? color2 + b * N * 3 : nullptr;
            f.y1 = fused ? y1 + b * N * 3 : const_cast<float *>(imga + b * N * 3);
            f.y2 = fused ? y2 + b * N * 3 : const_cast<float *>(imgb + b * N * 3);
            f.var1 = vara + b * N;
            f.var2 = varb + b * N;
            f.alb = albedo + b * N * 3;
            f.nrm = normal + b * N * 3;
            f.beta = static_cast<float *>(ws) + b * Nb;
            f.outa = outa + b * Nr * 3;
            f.outb = outb + b * Nr * 3;
            f.H = H;
            f.W = W;
        }
        if (int e = hip_check(xdn::launch_variance(n, fr.data(), 3, fused, S(stream)), "variance")) return e;
        if (int e = hip_check(xdn::launch_regression(n, fr.data(), win, s, S(stream)), "regression")) return e;
        if (int e = hip_check(xdn::launch_recon(n, fr.data(), win, s, rs, S(stream)), "recon")) return e;
    }
    return XDN_OK;
}

int xdn_cross_regression(const float *imga, const float *imgb, const float *albedo, const float *normal, float *outa,
                         float *outb, float *out_vara, float *out_varb, void *workspace, size_t workspace_bytes,
                         int batch, int height, int width, int win_size, int sparsity, int recon_scale, void *stream)
{
    return cross_regression_impl(nullptr, nullptr, imga, imgb, albedo, normal, nullptr, nullptr, outa, outb, out_vara,
                                 out_varb, workspace, workspace_bytes, batch, height, width, win_size, sparsity,
                                 recon_scale, stream);
}

int xdn_cross_regression_from_color(const float *color, const float *color2, const float *albedo, const float *normal,
                                    float *y1, float *y2, float *outa, float *outb, float *out_vara, float *out_varb,
                                    void *workspace, size_t workspace_bytes, int batch, int height, int width,
                                    int win_size, int sparsity, void *stream)
{
    if (batch != 0 && (!color || !color2 || !y1 || !y2))
        return fail(XDN_EINVAL, "color, color2, y1 and y2 are required");
    return cross_regression_impl(color, color2, nullptr, nullptr, albedo, normal, y1, y2, outa, outb, out_vara,
                                 out_varb, workspace, workspace_bytes, batch, height, width, win_size, sparsity, 1,
                                 stream);
}

namespace {
// Workspace of the cross-regression backward, per image (every buffer 256-B aligned).
struct CrgLayout {
    size_t var1, var2, beta, gq, u, gcen, evar, bytes;
    CrgLayout(int H, int W, int s)
    {
        const size_t N = px(H, W), C = (size_t)(H / s) * (W / s);
        size_t o = 0;
        var1 = o; o += al256(N * 4);
        var2 = o; o += al256(N * 4);
        beta = o; o += al256(C * 36 * 4);
        gq = o; o += al256(N * 2 * 16);
        u = o; o += al256(C * 36 * 4);
        gcen = o; o += al256(C * 24 * 4);
        evar = o; o += al256(N * 6 * 4);
        bytes = o;
    }
};
}  // namespace

size_t xdn_cross_regression_grad_workspace_size(int batch, int height, int width, int sparsity)
{
    if (batch <= 0 || height <= 0 || width <= 0 || sparsity <= 0) return 0;
    return (size_t)batch * CrgLayout(height, width, sparsity).bytes;
}

int xdn_cross_regression_grad(const float *imga, const float *imgb, const float *albedo, const float *normal,
                              const float *grad_outa, const float *grad_outb, const float *grad_vara,
                              const float *grad_varb, float *grad_imga, float *grad_imgb, float *grad_albedo,
                              float *grad_normal, void *workspace, size_t workspace_bytes, int B, int H, int W,
                              int win, int s, void *stream)
{
    if (int e = check_shape(B, H, W)) return e;
    if (win < 1 || s < 1) return fail(XDN_EINVAL, "win_size and sparsity must be >= 1");
    if (H % s != 0 || W % s != 0) return fail(XDN_EINVAL, "Image height and width must be divisible by sparsity");
    if (H <= win / 2 || W <= win / 2) return fail(XDN_EINVAL, "Image height and width must exceed win_size/2");
    if (B == 0) return XDN_OK;  // empty batch: nothing to read or write
    if (!imga || !imgb || !albedo || !normal || !grad_outa || !grad_outb || !grad_imga || !grad_imgb ||
        !grad_albedo || !grad_normal)
        return fail(XDN_EINVAL, "cross_regression_grad: null tensor");
    const size_t need = xdn_cross_regression_grad_workspace_size(B, H, W, s);
    if (!workspace || workspace_bytes < need)
        return fail(XDN_EINVAL, "workspace too small: %zu < %zu", workspace_bytes, need);
    const CrgLayout lay(H, W, s);
    const size_t N = px(H, W);
    for (int b0 = 0; b0 < B; b0 += XDN_MAX_FRAMES / 2)
    {
        const int n = B - b0 < XDN_MAX_FRAMES / 2 ? B - b0 : XDN_MAX_FRAMES / 2;
        std::vector<xdn::CrFrame> fw(n);
        std::vector<xdn::CrgFrame> fr(n);
        for (int i = 0; i < n; ++i)
        {
            const size_t b = (size_t)(b0 + i);
            char *w = static_cast<char *>(workspace) + b * lay.bytes;
            xdn::CrFrame &f = fw[i];
            f = xdn::CrFrame{};
            f.y1 = const_cast<float *>(imga + b * N * 3);
            f.y2 = const_cast<float *>(imgb + b * N * 3);
            f.var1 = reinterpret_cast<float *>(w + lay.var1);
            f.var2 = reinterpret_cast<float *>(w + lay.var2);
            f.alb = albedo + b * N * 3;
            f.nrm = normal + b * N * 3;
            f.beta = reinterpret_cast<float *>(w + lay.beta);
            f.H = H;
            f.W = W;
            xdn::CrgFrame &g = fr[i];
            g.y1 = f.y1; g.y2 = f.y2; g.alb = f.alb; g.nrm = f.nrm;
            g.var1 = f.var1; g.var2 = f.var2; g.beta = f.beta;
            g.gouta = grad_outa + b * N * 3;
            g.goutb = grad_outb + b * N * 3;
            g.gvara = grad_vara ? grad_vara + b * N : nullptr;
            g.gvarb = grad_varb ? grad_varb + b * N : nullptr;
            g.gq = reinterpret_cast<float4 *>(w + lay.gq);
            g.u = reinterpret_cast<float *>(w + lay.u);
            g.gcen = reinterpret_cast<float *>(w + lay.gcen);
            g.evar = reinterpret_cast<float *>(w + lay.evar);
            g.gya = grad_imga + b * N * 3;
            g.gyb = grad_imgb + b * N * 3;
            g.galb = grad_albedo + b * N * 3;
            g.gnrm = grad_normal + b * N * 3;
            g.H = H;
            g.W = W;
        }
        if (int e = hip_check(xdn::launch_variance(n, fw.data(), 3, false, S(stream)), "variance")) return e;
        if (int e = hip_check(xdn::launch_regression(n, fw.data(), win, s, S(stream)), "regression")) return e;
        if (int e = hip_check(xdn::launch_cross_regression_grad(n, fr.data(), win, s, S(stream)),
                              "cross_regression_grad"))
            return e;
    }
    return XDN_OK;
}

static int sf_common_check(int B, int H, int W, int win)
{
    if (int e = check_shape(B, H, W)) return e;
    if (win < 1) return fail(XDN_EINVAL, "win_size must be >= 1");
    if (win / 2 < 1) return fail(XDN_EINVAL, "win_size must be >= 2");
    if (H <= win / 2 || W <= win / 2) return fail(XDN_EINVAL, "Image height and width must exceed win_size/2");
    if ((size_t)H * W >= kFilterMaxPx) return fail(XDN_EINVAL, "frames of fewer than 2^32 pixels (got %d x %d)", H, W);
    return XDN_OK;
}

int xdn_spatial_filter(const float *imga, const float *imgb, const float *albedo, const float *normal,
                       const float *band, float *out_acca, float *out_accb, float *out_wgtsuma, float *out_wgtsumb,
                       int batch, int height, int width, int win_size, void *stream)
{
    if (int e = sf_common_check(batch, height, width, win_size)) return e;
    if (batch == 0) return XDN_OK;  // empty batch: nothing to read or write
    if (!imga || !imgb || !albedo || !normal || !band || !out_acca || !out_accb || !out_wgtsuma || !out_wgtsumb)
        return fail(XDN_EINVAL, "spatial_filter: null tensor");
    const size_t N = px(height, width);
    return sf_run(
        batch, height, width, win_size,
        [&](xdn::SfFrame &f, size_t b) {
            f.ia = imga + b * N * 3; f.ib = imgb + b * N * 3; f.alb = albedo + b * N * 3; f.nrm = normal + b * N * 3;
            f.band = band + b * N * 5;
            f.accA = out_acca + b * N * 3; f.accB = out_accb + b * N * 3;
            f.wsA = out_wgtsuma + b * N; f.wsB = out_wgtsumb + b * N;
        },
        [&](int n, const xdn::SfFrame *fr) { return xdn::launch_sf_forward(n, fr, win_size, S(stream)); },
        "spatial_filter");
}

int xdn_spatial_filter_grad(const float *grad_acca, const float *grad_accb, const float *grad_wgtsuma,
                            const float *grad_wgtsumb, const float *imga, const float *imgb, const float *img_albedo,
                            const float *img_normal, const float *band, float *out_grad_band, int batch, int height,
                            int width, int win_size, void *stream)
{
    if (int e = sf_common_check(batch, height, width, win_size)) return e;
    if (batch == 0) return XDN_OK;  // empty batch: nothing to read or write
    if (!grad_acca || !grad_accb || !grad_wgtsuma || !grad_wgtsumb || !imga || !imgb || !img_albedo || !img_normal ||
        !band || !out_grad_band)
        return fail(XDN_EINVAL, "spatial_filter_grad: null tensor");
    const size_t N = px(height, width);
    return sf_run(
        batch, height, width, win_size,
        [&](xdn::SfFrame &f, size_t b) {
            f.ia = imga + b * N * 3; f.ib = imgb + b * N * 3; f.alb = img_albedo + b * N * 3;
            f.nrm = img_normal + b * N * 3; f.band = band + b * N * 5;
            f.gaccA = grad_acca + b * N * 3; f.gaccB = grad_accb + b * N * 3;
            f.gwA = grad_wgtsuma + b * N; f.gwB = grad_wgtsumb + b * N;
            f.gband = out_grad_band + b * N * 5;
        },
        [&](int n, const xdn::SfFrame *fr) { return xdn::launch_sf_grad_band(n, fr, win_size, S(stream)); },
        "spatial_filter_grad");
}

int xdn_spatial_filter_grad_inputs(const float *grad_acca, const float *grad_accb, const float *grad_wgtsuma,
                                   const float *grad_wgtsumb, const float *imga, const float *imgb,
                                   const float *img_albedo, const float *img_normal, const float *band,
                                   float *grad_imga, float *grad_imgb, float *grad_albedo, float *grad_normal,
                                   int batch, int height, int width, int win_size, void *stream)
{
    if (int e = sf_common_check(batch, height, width, win_size)) return e;
    if (win_size / 2 > 7) return fail(XDN_EINVAL, "input gradients support win_size <= 15");
    if (batch == 0) return XDN_OK;  // empty batch: nothing to read or write
    if (!grad_acca || !grad_accb || !grad_wgtsuma || !grad_wgtsumb || !imga || !imgb || !img_albedo || !img_normal ||
        !band)
        return fail(XDN_EINVAL, "spatial_filter_grad_inputs: null tensor");
    const size_t N = px(height, width);
    auto off = [](float *p, size_t o) { return p ? p + o : nullptr; };
    return sf_run(
        batch, height, width, win_size,
        [&](xdn::SfFrame &f, size_t b) {
            f.ia = imga + b * N * 3; f.ib = imgb + b * N * 3; f.alb = img_albedo + b * N * 3;
            f.nrm = img_normal + b * N * 3; f.band = band + b * N * 5;
            f.gaccA = grad_acca + b * N * 3; f.gaccB = grad_accb + b * N * 3;
            f.gwA = grad_wgtsuma + b * N; f.gwB = grad_wgtsumb + b * N;
            f.gia = off(grad_imga, b * N * 3); f.gib = off(grad_imgb, b * N * 3);
            f.galb = off(grad_albedo, b * N * 3); f.gnrm = off(grad_normal, b * N * 3);
        },
        [&](int n, const xdn::SfFrame *fr) { return xdn::launch_sf_grad_inputs(n, fr, win_size, S(stream)); },
        "spatial_filter_grad_inputs");
}

int xdn_reproject(const float *mvec, int n_list, const float *const *inputs, const int *dims, const float *linear_z,
                  const float *prev_linear_z, const float *normal, const float *prev_normal, const float *pn_fwidth,
                  const float *opacity, const float *prev_opacity, uint8_t *success, float *const *outputs, int batch,
                  int height, int width, void *stream)
{
    if (int e = check_shape(batch, height, width)) return e;
    if (n_list < 0) return fail(XDN_EINVAL, "n_list must be >= 0");
    if (n_list > 0 && !dims) return fail(XDN_EINVAL, "reproject: null dims");
    for (int l = 0; l < n_list; ++l)
        if (dims[l] < 1 || dims[l] > 4) return fail(XDN_EINVAL, "input_list[%d] has %d channels (1..4 supported)", l, dims[l]);
    if (batch == 0) return XDN_OK;  // empty batch: nothing to read or write
    if (!mvec || !linear_z || !prev_linear_z || !normal || !prev_normal || !pn_fwidth || !opacity || !prev_opacity ||
        !success || (n_list > 0 && (!inputs || !outputs)))
        return fail(XDN_EINVAL, "reproject: null tensor");
    for (int l = 0; l < n_list; ++l)
        if (!inputs[l] || !outputs[l]) return fail(XDN_EINVAL, "reproject: input_list[%d] is null", l);
    const size_t N = px(height, width);
    // buffers beyond XDN_MAX_LIST go to further launches (the mask is recomputed, identically)
    int l0 = 0;
    do
    {
        const int nl = n_list - l0 < XDN_MAX_LIST ? n_list - l0 : XDN_MAX_LIST;
        for (int b0 = 0; b0 < batch; b0 += XDN_MAX_FRAMES)
        {
            const int n = batch - b0 < XDN_MAX_FRAMES ? batch - b0 : XDN_MAX_FRAMES;
            std::vector<xdn::RpFrame> fr(n);
            for (int i = 0; i < n; ++i)
            {
                const size_t b = (size_t)(b0 + i);
                xdn::RpFrame &f = fr[i];
                memset(&f, 0, sizeof(f));
                f.mvec = mvec + b * N * 3; f.linZ = linear_z + b * N * 3; f.prevLinZ = prev_linear_z + b * N * 3;
                f.nrm = normal + b * N * 3; f.prevNrm = prev_normal + b * N * 3; f.pnFw = pn_fwidth + b * N * 3;
                f.opac = opacity + b * N; f.prevOpac = prev_opacity + b * N; f.success = success + b * N;
                for (int l = 0; l < nl; ++l)
                {
                    f.in[l] = inputs[l0 + l] + b * N * dims[l0 + l];
                    f.out[l] = outputs[l0 + l] + b * N * dims[l0 + l];
                }
                f.H = height;
                f.W = width;
            }
            if (int e = hip_check(xdn::launch_reproject(n, fr.data(), nl, dims + l0, false, S(stream)), "reproject")) return e;
        }
        l0 += nl;
    } while (l0 < n_list);
    return XDN_OK;
}

static int screening_impl(const float *c1, const float *c2, const float *prev, const float *prev1, const float *prev2,
                          const float *rsucc, const uint8_t *mask, float *eroded, float *o, float *o1, float *o2,
                          float *osucc, int B, int H, int W, int win_curr, int win_prev, void *stream)
{
    if (int e = check_shape(B, H, W)) return e;
    if (win_curr < 1 || win_prev < 1) return fail(XDN_EINVAL, "win_curr and win_prev must be >= 1");
    if (B == 0) return XDN_OK;  // empty batch: nothing to read or write
    if (!c1 || !c2 || !prev || !prev1 || !prev2 || !o || !o1 || !o2 || !osucc)
        return fail(XDN_EINVAL, "temporal_screening: null tensor");
    const size_t N = px(H, W);
    for (int b0 = 0; b0 < B; b0 += XDN_MAX_FRAMES)
    {
        const int n = B - b0 < XDN_MAX_FRAMES ? B - b0 : XDN_MAX_FRAMES;
        std::vector<xdn::TsFrame> fr(n);
        for (int i = 0; i < n; ++i)
        {
            const size_t b = (size_t)(b0 + i);
            xdn::TsFrame &f = fr[i];
            f.c1 = c1 + b * N * 3; f.c2 = c2 + b * N * 3; f.prev = prev + b * N * 3;
            f.prev1 = prev1 + b * N * 3; f.prev2 = prev2 + b * N * 3;
            f.rsucc = rsucc ? rsucc + b * N : nullptr;
            f.succ_mask = mask ? mask + b * N : nullptr;
            f.succ_e = eroded ? eroded + b * N : nullptr;
            f.o = o + b * N * 3; f.o1 = o1 + b * N * 3; f.o2 = o2 + b * N * 3; f.osucc = osucc + b * N;
            f.H = H;
            f.W = W;
        }
        if (int e = hip_check(xdn::launch_screening(n, fr.data(), win_curr, win_prev, mask != nullptr, S(stream)),
                              "temporal_screening"))
            return e;
    }
    return XDN_OK;
}

int xdn_temporal_screening(const float *current1, const float *current2, const float *prev, const float *prev1,
                           const float *prev2, const float *reproj_success, float *output, float *output1,
                           float *output2, float *success, int batch, int height, int width, int win_curr,
                           int win_prev, void *stream)
{
    if (batch != 0 && !reproj_success) return fail(XDN_EINVAL, "reproj_success is required");
    return screening_impl(current1, current2, prev, prev1, prev2, reproj_success, nullptr, nullptr, output, output1,
                          output2, success, batch, height, width, win_curr, win_prev, stream);
}

int xdn_erode_temporal_screening(const float *current1, const float *current2, const float *prev, const float *prev1,
                                 const float *prev2, const uint8_t *success_mask, float *eroded, float *output,
                                 float *output1, float *output2, float *success, int batch, int height, int width,
                                 int win_curr, int win_prev, void *stream)
{
    if (batch != 0 && !success_mask) return fail(XDN_EINVAL, "success_mask is required");
    return screening_impl(current1, current2, prev, prev1, prev2, nullptr, success_mask, eroded, output, output1,
                          output2, success, batch, height, width, win_curr, win_prev, stream);
}

size_t xdn_filter_train_workspace_size(int batch, int height, int width)
{
    if (batch <= 0 || height <= 0 || width <= 0) return 0;
    return (size_t)batch * (size_t)xdn::filter_train_tiles(height, width) * sizeof(float);
}

int xdn_filter_train(const float *ols1, const float *ols2, const float *albedo, const float *normal, const float *param,
                     const float *success, const float *success_prev, const float *prev_out, const float *opacity,
                     const float *prev_ols1, const float *prev_ols2, float *out_log, float *out, float *dparam,
                     float *loss, float *avg_loss, void *workspace, size_t workspace_bytes, int batch, int height,
                     int width, int win_size, void *stream)
{
    if (int e = sf_common_check(batch, height, width, win_size)) return e;
    if (win_size / 2 != 5) return fail(XDN_EINVAL, "the fused training kernel supports win_size 11 (half 5) only");
    if (batch == 0) return XDN_OK;  // empty batch: nothing to read or write
    if (!ols1 || !ols2 || !albedo || !normal || !param || !success || !success_prev || !prev_out || !opacity ||
        !prev_ols1 || !prev_ols2 || !out_log || !out || !dparam)
        return fail(XDN_EINVAL, "filter_train: null tensor");
    if (!workspace || workspace_bytes < xdn_filter_train_workspace_size(batch, height, width))
        return fail(XDN_EINVAL, "workspace too small");
    const size_t N = px(height, width);
    const int T = xdn::filter_train_tiles(height, width);
    float *parts = static_cast<float *>(workspace);
    for (int b0 = 0; b0 < batch; b0 += XDN_MAX_FRAMES)
    {
        const int n = batch - b0 < XDN_MAX_FRAMES ? batch - b0 : XDN_MAX_FRAMES;
        std::vector<xdn::TrainFrame> fr(n);
        std::vector<const float *> pp(n);
        std::vector<int> cnt(n);
        std::vector<float> sc(n);
        std::vector<float *> outp(n);
        for (int i = 0; i < n; ++i)
        {
            const size_t b = (size_t)(b0 + i);
            xdn::TrainFrame &f = fr[i];
            f.ols1 = ols1 + b * N * 3; f.ols2 = ols2 + b * N * 3; f.alb = albedo + b * N * 3; f.nrm = normal + b * N * 3;
            f.param = param + b * N * 6; f.succ_e = success + b * N; f.succ_prev = success_prev + b * N;
            f.prev_out = prev_out + b * N * 3; f.opacity = opacity + b * N;
            f.prev_ols1 = prev_ols1 + b * N * 3; f.prev_ols2 = prev_ols2 + b * N * 3;
            f.out_log = out_log + b * N * 3; f.out = out + b * N * 3; f.dparam = dparam + b * N * 6;
            f.loss = loss ? loss + b * N * 3 : nullptr;
            f.loss_part = parts + b * T;
            f.H = height;
            f.W = width;
            pp[i] = f.loss_part;
            cnt[i] = T;
            sc[i] = (float)(1.0 / (3.0 * (double)N));
            outp[i] = avg_loss + b;
        }
        if (int e = hip_check(xdn::launch_filter_train(n, fr.data(), win_size, S(stream)), "filter_train")) return e;
        if (avg_loss)
            if (int e = hip_check(xdn::launch_sums(n, pp.data(), cnt.data(), sc.data(), outp.data(), S(stream)), "loss sum"))
                return e;
    }
    return XDN_OK;
}

size_t xdn_frame_error_workspace_size(int batch, int height, int width)
{
    if (batch <= 0 || height <= 0 || width <= 0) return 0;
    return (size_t)batch * xdn::frame_error_blocks(height, width) * sizeof(float);
}

int xdn_frame_error(const float *out, const float *emissive, const float *env_light, const float *ref,
                    const float *ref_emissive, const float *ref_env_light, float *out_composed, float *err,
                    void *workspace, size_t workspace_bytes, int batch, int height, int width, void *stream)
{
    if (int e = check_shape(batch, height, width)) return e;
    if (batch == 0) return XDN_OK;  // empty batch: nothing to read or write
    if (!out || !ref || !err) return fail(XDN_EINVAL, "frame error: out, ref and err are required");
    if (!workspace || workspace_bytes < xdn_frame_error_workspace_size(batch, height, width))
        return fail(XDN_EINVAL, "frame error: workspace too small");
    const size_t N3 = px(height, width) * 3;
    const int nb = xdn::frame_error_blocks(height, width);
    float *parts = static_cast<float *>(workspace);
    auto at = [&](const float *p, int b) { return p ? p + (size_t)b * N3 : nullptr; };
    for (int b0 = 0; b0 < batch; b0 += XDN_MAX_FRAMES)
    {
        const int n = batch - b0 < XDN_MAX_FRAMES ? batch - b0 : XDN_MAX_FRAMES;
        xdn::ErrFrame fr[XDN_MAX_FRAMES];
        const float *pp[XDN_MAX_FRAMES];
        int cnt[XDN_MAX_FRAMES];
        float scale[XDN_MAX_FRAMES];
        float *dst[XDN_MAX_FRAMES];
        for (int i = 0; i < n; ++i)
        {
            const int b = b0 + i;
            fr[i] = xdn::ErrFrame{at(out, b), at(emissive, b), at(env_light, b), at(ref, b), at(ref_emissive, b),
                                  at(ref_env_light, b), out_composed ? out_composed + (size_t)b * N3 : nullptr,
                                  parts + (size_t)b * nb, height, width};
            pp[i] = parts + (size_t)b * nb;
            cnt[i] = nb;
            scale[i] = (float)(1.0 / (double)N3);
            dst[i] = err + b;
        }
        if (int e = hip_check(xdn::launch_frame_error(n, fr, S(stream)), "frame error")) return e;
        if (int e = hip_check(xdn::launch_sums(n, pp, cnt, scale, dst, S(stream)), "frame error sum")) return e;
    }
    return XDN_OK;
}

int xdn_adam(float *param, const float *grad, float *m, float *v, int count, int step, float lr, float beta1,
             float beta2, float eps, void *stream)
{
    if (count < 0 || step < 1) return fail(XDN_EINVAL, "count must be >= 0 and step >= 1");
    if (count > 0 && (!param || !grad || !m || !v)) return fail(XDN_EINVAL, "adam: null tensor");
    xdn::AdamJob j{};
    j.param = param;
    j.grad = const_cast<float *>(grad);
    j.m = m;
    j.v = v;
    j.count = count;
    j.alpha = (float)((double)lr * sqrt(1.0 - pow((double)beta2, step)) / (1.0 - pow((double)beta1, step)));
    return hip_check(xdn::launch_adam(1, &j, beta1, beta2, eps, S(stream)), "adam");
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// cnn5: the bandwidth network (5 conv3x3: 15 -> 32 -> 32 -> 32 -> 32 -> 6)
// ---------------------------------------------------------------------------------------------

extern "C" {

size_t xdn_cnn5_param_count(void) { return kLay.count; }

size_t xdn_cnn5_packed_bytes(void) { return kLay.packed; }

size_t xdn_cnn5_workspace_size(int batch, int height, int width)
{
    if (batch <= 0 || height <= 0 || width <= 0) return 0;
    return Cnn5Ws(nullptr, batch, height, width).bytes;
}

int xdn_cnn5_pack(const float *params, void *packed, void *stream)
{
    if (!params || !packed) return fail(XDN_EINVAL, "cnn5 pack: null buffer");
    xdn::PackJob j[5];
    char *pk = static_cast<char *>(packed);
    for (int l = 0; l < 5; ++l)
    {
        j[l].w = params + kLay.w_off[l];
        j[l].b = params + kLay.b_off[l];
        j[l].frag_fwd = pk + kLay.fwd_off[l];
        j[l].frag_dgrad = l > 0 ? pk + kLay.dg_off[l] : nullptr;
        j[l].bias32 = reinterpret_cast<float *>(pk + kLay.bias_off[l]);
        j[l].ci = kCnn5[l].ci;
        j[l].co = kCnn5[l].co;
    }
    return hip_check(xdn::launch_pack(5, j, S(stream)), "cnn5 pack");
}

static int cnn5_check(int B, int H, int W, size_t ws_bytes)
{
    if (int e = check_shape(B, H, W)) return e;
    if (B > XDN_MAX_FRAMES) return fail(XDN_EINVAL, "cnn5: batch must be <= %d", XDN_MAX_FRAMES);
    if ((size_t)H * W >= kCnn5MaxPx) return fail(XDN_EINVAL, "cnn5: frames of fewer than 2^25 pixels (got %d x %d)", H, W);
    if (ws_bytes < xdn_cnn5_workspace_size(B, H, W)) return fail(XDN_EINVAL, "cnn5: workspace too small");
    return XDN_OK;
}

int xdn_cnn5_forward(const float *ols1, const float *ols2, const float *prev_out, const float *albedo,
                     const float *normal, const void *packed, float *param, void *workspace, size_t workspace_bytes,
                     int batch, int height, int width, void *stream)
{
    if (int e = cnn5_check(batch, height, width, workspace_bytes)) return e;
    if (batch == 0) return XDN_OK;  // empty batch: nothing to read or write
    if (!ols1 || !ols2 || !prev_out || !albedo || !normal || !packed || !param || !workspace)
        return fail(XDN_EINVAL, "cnn5 forward: null buffer");
    Cnn5Ws ws(workspace, batch, height, width);
    const size_t N = px(height, width);
    const char *pk = static_cast<const char *>(packed);
    // layer 1 on its own (the five FP32 images in), layers 2..5 in one row-streaming launch
    std::vector<xdn::ConvFrame> fr(batch);
    std::vector<xdn::FwdStreamFrame> sf(batch);
    for (int b = 0; b < batch; ++b)
    {
        xdn::ConvFrame &f = fr[b];
        memset(&f, 0, sizeof(f));
        const size_t o3 = (size_t)b * N * 3, o32 = (size_t)b * N * 32;
        f.in[0] = ols1 + o3; f.in[1] = ols2 + o3; f.in[2] = prev_out + o3; f.in[3] = albedo + o3;
        f.in[4] = normal + o3;
        f.out = ws.X[0] + o32;
        f.wfrag = pk + kLay.fwd_off[0];
        f.bias = reinterpret_cast<const float *>(pk + kLay.bias_off[0]);
        f.H = height;
        f.W = width;
        sf[b] = cnn5_stream_frame(ws, pk, (size_t)b * N, param + (size_t)b * N * 6, height, width);
    }
    if (int e = hip_check(xdn::launch_conv_l1(batch, fr.data(), S(stream)), "cnn5 forward")) return e;
    return hip_check(xdn::launch_fwd_stream(batch, sf.data(), S(stream)), "cnn5 forward");
}

int xdn_cnn5_backward(const float *ols1, const float *ols2, const float *prev_out, const float *albedo,
                      const float *normal, const float *dparam, const void *packed, float *grad, void *workspace,
                      size_t workspace_bytes, int batch, int height, int width, void *stream)
{
    if (int e = cnn5_check(batch, height, width, workspace_bytes)) return e;
    if (batch == 0)  // empty batch: the gradient is a sum over no images
    {
        if (!grad) return fail(XDN_EINVAL, "cnn5 backward: null buffer");
        return hip_check(xdn::launch_zero(grad, kLay.count, S(stream)), "cnn5 backward");
    }
    if (!ols1 || !ols2 || !prev_out || !albedo || !normal || !dparam || !packed || !grad || !workspace)
        return fail(XDN_EINVAL, "cnn5 backward: null buffer");
    Cnn5Ws ws(workspace, batch, height, width);
    const size_t N = px(height, width);
    const char *pk = static_cast<const char *>(packed);
    std::vector<xdn::RedJob> jobs;
    std::vector<xdn::BwdFrame> bf(batch);
    int cur = 0;  // dY ping-pong index holding the gradient of the current layer's output
    // layers 5..3: weight gradient and data gradient into layer l-1 in one kernel (the last layer's
    // dY is the FP32 gradient of the network output); the B images' partials of a layer are reduced
    // together into the one gradient
    for (int l = 4; l >= 2; --l)
    {
        const int nb = ws.nblk[l];
        for (int b = 0; b < batch; ++b)
        {
            const size_t o32 = (size_t)b * N * 32;
            const void *dy = l == 4 ? static_cast<const void *>(dparam + (size_t)b * N * 6)
                                    : static_cast<const void *>(ws.dY[cur] + o32);
            bf[b] = xdn::BwdFrame{dy, ws.X[l - 1] + o32, ws.dY[1 - cur] + o32, pk + kLay.dg_off[l],
                                  ws.part[l] + (size_t)b * nb * ws.pstride[l], height, width, nb};
        }
        if (int e = hip_check(xdn::launch_bwd(batch, bf.data(), l == 4, S(stream)), "cnn5 backward")) return e;
        jobs.push_back(cnn5_red_job(ws, l, grad, batch * nb));
        cur = 1 - cur;
    }
    // layer 2's backward with layer 1's weight gradient (x = the five input images; layer 1's
    // pre-activation gradient stays on chip)
    std::vector<xdn::Bwd12Frame> f12(batch);
    const int nb1 = ws.nblk[1], nb0 = ws.nblk[0];
    for (int b = 0; b < batch; ++b)
    {
        const size_t o3 = (size_t)b * N * 3, o32 = (size_t)b * N * 32;
        f12[b] = xdn::Bwd12Frame{ws.dY[cur] + o32, ws.X[0] + o32, pk + kLay.dg_off[1],
                                 ws.part[1] + (size_t)b * nb1 * ws.pstride[1],
                                 {ols1 + o3, ols2 + o3, prev_out + o3, albedo + o3, normal + o3},
                                 ws.part[0] + (size_t)b * nb0 * ws.pstride[0], height, width, nb1};
    }
    if (int e = hip_check(xdn::launch_bwd12(batch, f12.data(), S(stream)), "cnn5 backward")) return e;
    jobs.push_back(cnn5_red_job(ws, 1, grad, batch * nb1));
    jobs.push_back(cnn5_red_job(ws, 0, grad, batch * nb0));
    return hip_check(xdn::launch_wgrad_reduce((int)jobs.size(), jobs.data(), S(stream)), "cnn5 wgrad reduce");
}

}  // extern "C"
