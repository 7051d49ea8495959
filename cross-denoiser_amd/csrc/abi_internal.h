// Internal helpers shared by the C-ABI translation units of libxdn.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "../../include/xdn.h"
#include "xdn_cnn.h"
#include "xdn_common.h"

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
// the weight-gradient partials of every block and the reduction stage (all five layers are
// reduced together after the backward pass).
struct Cnn5Ws {
    uint16_t *X[4], *dY[2];
    float *part[5], *stage[5];
    int nblk[5];  // weight-gradient blocks per image and layer
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
            nblk[l] = xdn::wgrad_blocks(H, W);
            part[l] = reinterpret_cast<float *>(p + o);
            o += al256((size_t)B * nblk[l] * xdn::WG_PART * 4);
            stage[l] = reinterpret_cast<float *>(p + o);
            o += al256((size_t)xdn::WG_SPLITS * xdn::WG_PART * 4);
        }
        bytes = o;
    }
};

}  // namespace xdn_abi
