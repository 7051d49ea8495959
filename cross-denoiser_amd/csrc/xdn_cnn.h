// Launchers of the MFMA convolution kernels of the bandwidth CNN (internal header).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xdn_common.h"

namespace xdn {

enum { CONV_IN_BF16 = 0, CONV_IN_IMAGES = 1, CONV_IN_F32_6 = 2 };
enum { CONV_EPI_RELU = 0, CONV_EPI_MASK = 1, CONV_EPI_F32_6 = 2 };
enum {
    CONV_L1_FWD = 0,    // 5 FP32 3-ch images (15 ch) -> bf16 32 ch, bias + ReLU
    CONV_MID_FWD = 1,   // bf16 32 -> bf16 32, bias + ReLU
    CONV_L5_FWD = 2,    // bf16 32 -> FP32 6, bias
    CONV_MID_DGRAD = 3, // bf16 32 (dY) -> bf16 32, ReLU mask of the layer input
    CONV_L5_DGRAD = 4,  // FP32 6 (dY) -> bf16 32, ReLU mask
};
enum { CONV_L1_WGRAD = 0, CONV_MID_WGRAD = 1, CONV_L5_WGRAD = 2 };

constexpr int WG_PART = 9 * 1024 + 32;  // one block's partial: [9][32][32] weights + [32] bias
constexpr int WG_SPLITS = 32;

struct ConvFrame {
    const void *in[5];       // see the kernel kinds
    void *out;
    const uint32_t *mask;    // [H][W] ReLU bits of the layer input, bit c = channel c > 0 (mask epilogue)
    uint32_t *mask_out;      // [H][W] ReLU bits of the output (bias + ReLU epilogue; may be null)
    const void *wfrag;       // 16-byte B fragments [9][KC][64]
    const float *bias;       // [32]
    int H, W;
};

struct WgFrame {
    const void *x[5];        // bf16 [H][W][32] activation, or the 5 FP32 input images
    const void *dy;          // bf16 [H][W][32] or FP32 [H][W][6]
    float *part;             // [nblk][WG_PART]
    int H, W, nblk;
};

// Fused backward of one 32 -> 32 layer: its weight gradient and its data gradient into the layer
// below, both from one staging of dY.
struct BwdFrame {
    const void *dy;          // bf16 [H][W][32] (last layer: FP32 [H][W][6]): gradient of the layer's output
    const void *x;           // bf16 [H][W][32]: the layer's input
    const uint32_t *mask;    // [H][W] ReLU bits of x
    void *dx;                // bf16 [H][W][32]: gradient of x, ReLU-masked (dY of the layer below)
    const void *wfrag;       // data-gradient B fragments [9][KC][64] (16 B each; KC 2, last layer 1)
    float *part;             // [nblk][WG_PART] weight-gradient partials
    int H, W, nblk;
};

// Layers 2..5 of the cnn5 forward in one row-streaming launch: a block walks a 26-px-wide strip of
// rows down a segment of the frame, each wave computing one layer's row per step from the rows
// of the layer below held in LDS.
struct FwdStreamFrame {
    const void *x1;          // bf16 [H][W][32]: output of layer 1
    void *x[3];              // bf16 [H][W][32]: outputs of layers 2, 3, 4
    uint32_t *m[3];          // their ReLU bits
    float *out;              // FP32 [H][W][6]: network output (layer 5)
    const void *wfrag[4];    // forward B fragments [9][2][64] (16 B each) of layers 2..5
    const float *bias[4];    // [32] bias of layers 2..5 (zero padded)
    int H, W;
};

struct PackJob {
    const float *w, *b;      // FP32 master [9][ci][co], [co]
    void *frag_fwd, *frag_dgrad;  // dgrad may be null
    float *bias32;
    int ci, co;
};

hipError_t launch_conv(int kind, int n, const ConvFrame *fr, hipStream_t st);
hipError_t launch_wgrad(int kind, int n, const WgFrame *fr, hipStream_t st);
// One reduction job: the nblk partials of one layer of one frame -> its flat gradient.
// The partials hold an input-channel slice ci_off .. ci_off + ci_cnt - 1 (ci_cnt < 0: all ci),
// [taps][32 ci][32 co] + [32 co] in WG_PART floats each, or when compact [taps][ci_cnt][co] +
// [co] in pstride floats each; the gradient is written in the layer's flat layout
// [taps][ci][co] (or [co][ci] when transpose, taps 1) followed by [co] when bias.
struct RedJob {
    const float *part;  // [nblk][pstride]
    float *stage;       // [WG_SPLITS][pstride]
    float *grad;        // the layer's weights, then its bias
    int nblk, ci, co;
    int taps = 9, ci_off = 0, ci_cnt = -1, bias = 1, transpose = 0;
    int compact = 0, pstride = WG_PART;
};
// floats of one compact partial: [taps][cc][co] + [co]
constexpr int compact_part(int taps, int cc, int co) { return taps * cc * co + co; }
constexpr int WG_MAX_JOBS = 48;  // per launch; longer job lists are split
hipError_t launch_wgrad_reduce(int n, const RedJob *jobs, hipStream_t st);
hipError_t launch_pack(int n, const PackJob *jobs, hipStream_t st);
int wgrad_blocks(int H, int W);
// last: the network's last layer (dy = the FP32 6-channel gradient of the output, L5 fragments)
hipError_t launch_bwd(int n, const BwdFrame *fr, bool last, hipStream_t st);
hipError_t launch_fwd_stream(int n, const FwdStreamFrame *fr, hipStream_t st);
int bwd_blocks(int H, int W);  // blocks (= weight-gradient partials) per image of launch_bwd

}  // namespace xdn
