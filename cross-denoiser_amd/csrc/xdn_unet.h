// The reference's bandwidth U-Net (scripts/main.py:83-118) on gfx950: frame descriptions and
// launchers of its kernels (internal header).
//
// Network (Keras order; H0 = full resolution, H1 = H0/2, H2 = H0/4; activations bf16 NHWC):
//   0 c1_1  conv3x3 15->8  ReLU  H0      (input: the five 3-channel FP32 images)
//   1 c1_2  conv3x3  8->8  ReLU  H0
//     pool  2x2 max             H0->H1
//   2 c2_1  conv3x3  8->16 ReLU  H1
//   3 c2_2  conv3x3 16->16 ReLU  H1
//     pool  2x2 max             H1->H2
//   4 c3_1  conv3x3 16->32 ReLU  H2
//   5 c3_2  conv3x3 32->32 ReLU  H2
//   6 c3_3  convT 1x1 stride 2 32->32 (linear) H2->H1: out(2y,2x) = W^T in(y,x) + b, else b
//   7 c4_1  conv3x3 48->16 ReLU  H1      (input: concat[c3_3, c2_2])
//   8 c4_2  conv3x3 16->16 ReLU  H1
//   9 c4_3  convT 1x1 stride 2 16->16 H1->H0
//  10 c5_1  conv3x3 24->8  ReLU  H0      (input: concat[c4_3, c1_2])
//  11 head  conv1x1  8->6        H0      FP32 output (the band/alpha head follows in the filter)
// 31,414 parameters.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xdn_common.h"
#include "xdn_cnn.h"

namespace xdn {

enum { UC_SRC_BF16 = 0, UC_SRC_IMAGES = 1 };
enum { UC_EPI_RELU = 0, UC_EPI_MASK = 1, UC_EPI_PLAIN = 2 };

// 3x3 'same' convolution over concat[a (CA ch), b (CB ch)] (or the five images), forward or data
// gradient (the data gradient is the same kernel on dY with the flipped, transposed fragments).
struct UcFrame {
    const void *a, *b;          // bf16 sources (b may be null when CB == 0)
    const float *img[5];        // FP32 3-channel images (UC_SRC_IMAGES)
    void *out;                  // bf16 [H][W][CO]
    const uint16_t *mask;       // bf16 [H][W][CO] activation whose ReLU masks the output (MASK)
    const void *wfrag;          // 16-byte B fragments [9][KC][NT][64]
    const float *bias;          // [NT * 32] (RELU)
    const float *head_w;        // UC_F_24_8_HEAD: the 1x1 head, FP32 [8][6] then [6]
    float *head_out;            // UC_F_24_8_HEAD: its output, FP32 [H][W][6]
    int H, W;
};
// kind = index into the instantiated (SRC, CA, CB, CO, EPI) table of k_unet.hip
hipError_t launch_uconv(int kind, int n, const UcFrame *fr, hipStream_t st);
enum {
    UC_F_0 = 0,   // images -> 8, ReLU
    UC_F_8_8,     // 8 -> 8 ReLU
    UC_F_8_16,    // 8 -> 16 ReLU
    UC_F_16_16,   // 16 -> 16 ReLU
    UC_F_16_32,   // 16 -> 32 ReLU
    UC_F_32_32,   // 32 -> 32 ReLU
    UC_F_48_16,   // concat 32+16 -> 16 ReLU
    UC_F_24_8,    // concat 16+8 -> 8 ReLU
    UC_D_8_8,     // dgrad dY 8 -> dX 8, masked
    UC_D_16_8,    // dgrad dY 16 -> dX 8, unmasked (pool input)
    UC_D_16_16,   // dgrad dY 16 -> dX 16, masked
    UC_D_32_16,   // dgrad dY 32 -> dX 16, unmasked (pool input)
    UC_D_32_32,   // dgrad dY 32 -> dX 32, masked
    UC_D_16_48,   // dgrad dY 16 -> dX 48, unmasked (concat input)
    UC_D_8_24,    // dgrad dY 8 -> dX 24, unmasked (concat input)
    UC_F_24_8_HEAD,  // UC_F_24_8 with the 1x1 head 8 -> 6 (FP32) applied to its bf16 output
    UC_KINDS
};

// Weight gradient of a 3x3 layer over concat[a (CA), b (CB)] (or the images) and dY (CO ch): per
// block one compact partial [9][cs][CO] + [CO] of the input slice [32 * mtile, 32 * mtile + cs)
// (cs = min(32, channels - 32 * mtile); compact_part(9, cs, CO) floats).
struct UwFrame {
    const void *a, *b;
    const float *img[5];
    const uint16_t *dy;         // bf16 [H][W][CO]
    float *part;                // [nblk][compact_part(9, cs, CO)]
    int H, W, nblk;
};
enum {
    UW_0 = 0,     // images (15) x dY 8
    UW_8_8,       // 8 x 8
    UW_8_16,      // 8 x 16
    UW_16_16,     // 16 x 16
    UW_16_32,     // 16 x 32
    UW_32_32,     // 32 x 32
    UW_48_16_M0,  // concat 32+16 x 16, input channels 0-31
    UW_48_16_M1,  // concat 32+16 x 16, input channels 32-47
    UW_24_8,      // concat 16+8 x 8
    UW_KINDS
};
hipError_t launch_uwgrad(int kind, int n, const UwFrame *fr, hipStream_t st);

// 2x2 max pool (valid), bf16 [H][W][C] -> [H/2][W/2][C].
struct PoolFrame {
    const uint16_t *in;
    uint16_t *out;
    int H, W;  // of the input
};
hipError_t launch_maxpool(int C, int n, const PoolFrame *fr, hipStream_t st);

// Gradient into a pooled activation act [H][W][C]: the pooled gradient (bf16 [H/2][W/2][C]) goes to
// the first maximum of each 2x2 window (row-major), the skip gradient (bf16 [H][W][sstride], channels
// soff..soff+C-1) is added in FP32, and the ReLU of act masks the sum -> bf16 dY [H][W][C].
struct PoolBwFrame {
    const uint16_t *dpool;
    const uint16_t *act;
    const uint16_t *dskip;
    uint16_t *dy;
    int H, W, sstride, soff;
};
hipError_t launch_pool_bwd(int C, int n, const PoolBwFrame *fr, hipStream_t st);

// Transposed 1x1 convolution, stride 2 ('same'): out(2y,2x) = W^T in(y,x) + b, other positions b.
// W is the Keras kernel [co][ci] (shape 1x1xCOxCI). Backward: dIn(y,x) = W dOut(2y,2x) masked by
// the ReLU of in -> bf16; per block one compact partial: dW [ci][co] then db (all output positions).
struct CtFrame {
    const uint16_t *in;         // bf16 [h][w][C]
    uint16_t *out;              // bf16 [2h][2w][C] (forward)
    const uint16_t *dout;       // bf16 [2h][2w][dstride], channels 0..C-1 (backward)
    uint16_t *din;              // bf16 [h][w][C] (backward)
    const float *w, *b;         // FP32 master
    float *part;                // [nblk][C * C + C] (backward)
    int h, w_, dstride, nblk;
};
hipError_t launch_convT(int C, int n, const CtFrame *fr, hipStream_t st);
hipError_t launch_convT_bwd(int C, int n, const CtFrame *fr, hipStream_t st);

// Head (1x1 convolution 8 -> 6, FP32 output; its forward is UC_F_24_8_HEAD) backward: d(in) =
// W dparam masked by the ReLU of in (bf16), per block one compact partial [8][6] + [6].
struct HeadFrame {
    const uint16_t *in;         // bf16 [H][W][8]
    const float *dparam;        // FP32 [H][W][6]
    uint16_t *din;              // bf16 [H][W][8]
    const float *w, *b;         // FP32 master [8][6], [6]
    float *part;
    int H, W, nblk;
};
hipError_t launch_head_bwd(int n, const HeadFrame *fr, hipStream_t st);

// Weight packing of one 3x3 layer: forward fragments [9][KC][NT][64] (K = ci, N = co) and, when
// frag_dgrad is set, data-gradient fragments (K = co, N = ci, taps flipped). A job with copy_n > 0
// instead copies copy_n floats from w to frag_fwd (the FP32 1x1 layers).
struct UPackJob {
    const float *w, *b;         // FP32 master [3][3][ci][co], [co]
    void *frag_fwd, *frag_dgrad;
    float *bias32;              // [NT * 32]
    int ci, co, copy_n;
};
hipError_t launch_upack(int n, const UPackJob *jobs, hipStream_t st);

int ublocks(int H, int W);      // weight-gradient blocks per image

}  // namespace xdn
