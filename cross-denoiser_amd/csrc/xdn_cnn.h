// Launchers of the MFMA convolution kernels of the bandwidth CNN (internal header).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xdn_common.h"

namespace xdn {

constexpr int WG_PART = 9 * 1024 + 32;  // one block's partial: [9][32][32] weights + [32] bias
constexpr int WG_SPLITS = 32;

// Layer 1 forward: the five FP32 3-channel images -> bf16 32 channels, bias + ReLU.
struct ConvFrame {
    const float *in[5];      // FP32 [H][W][3] each: ols1, ols2, prev_out, albedo, normal
    void *out;               // bf16 [H][W][32]
    const void *wfrag;       // 16-byte B fragments [9][64]
    const float *bias;       // [32]
    int H, W;
};

// Fused backward of one 32 -> 32 layer: its weight gradient and its data gradient into the layer
// below, both from one staging of dY and x.
struct BwdFrame {
    const void *dy;          // bf16 [H][W][32] (last layer: FP32 [H][W][6]): gradient of the layer's output
    const void *x;           // bf16 [H][W][32]: the layer's input (a ReLU output)
    void *dx;                // bf16 [H][W][32]: gradient of x where x > 0, else 0 (dY of the layer below)
    const void *wfrag;       // data-gradient B fragments [9][KC][64] (16 B each; KC 2, last layer 1)
    float *part;             // [nblk][WG_PART] weight-gradient partials (last layer: WG_PART_L5, compact)
    int H, W, nblk;
};

// Backward of layer 2 fused with layer 1's weight gradient: layer 2's weight gradient and its data
// gradient, which (masked by x > 0) is layer 1's pre-activation gradient; that one stays in LDS and
// meets the five input images there for layer 1's weight and bias gradient, so it is never written.
struct Bwd12Frame {
    const void *dy;          // bf16 [H][W][32]: gradient of layer 2's pre-activation
    const void *x;           // bf16 [H][W][32]: layer 2's input (layer 1's ReLU output)
    const void *wfrag;       // layer 2's data-gradient B fragments [9][2][64]
    float *part;             // [nblk][WG_PART] layer 2's weight-gradient partials
    const float *img[5];     // FP32 [H][W][3]: layer 1's input images
    float *part1;            // [nblk][WG_PART_L1] layer 1's weight-gradient partials, compact
    int H, W, nblk;
};

// Layers 2..5 of the cnn5 forward in one row-streaming launch: a block walks a 26-px-wide strip of
// rows down a segment of the frame, each wave computing one layer's row per step from the rows
// of the layer below held in LDS.
struct FwdStreamFrame {
    const void *x1;          // bf16 [H][W][32]: output of layer 1
    void *x[3];              // bf16 [H][W][32]: outputs of layers 2, 3, 4
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

hipError_t launch_conv_l1(int n, const ConvFrame *fr, hipStream_t st);
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
// the first layer's (15 -> 32) and the last layer's (32 -> 6) partials are compact
constexpr int WG_PART_L1 = compact_part(9, 15, 32);
constexpr int WG_PART_L5 = compact_part(9, 32, 6);
constexpr int WG_MAX_JOBS = 48;  // per launch; longer job lists are split
hipError_t launch_wgrad_reduce(int n, const RedJob *jobs, hipStream_t st);
hipError_t launch_pack(int n, const PackJob *jobs, hipStream_t st);
// last: the network's last layer (dy = the FP32 6-channel gradient of the output, L5 fragments)
hipError_t launch_bwd(int n, const BwdFrame *fr, bool last, hipStream_t st);
hipError_t launch_bwd12(int n, const Bwd12Frame *fr, hipStream_t st);  // nblk = bwd_blocks(H, W)
hipError_t launch_fwd_stream(int n, const FwdStreamFrame *fr, hipStream_t st);
int bwd_blocks(int H, int W);  // blocks (= weight-gradient partials) per image of launch_bwd

}  // namespace xdn
