// Host-side frame descriptions and kernel launchers of the xdn library (internal header).
//
// A launcher takes n frames (n <= XDN_MAX_FRAMES unless stated), each with its own H, W and
// buffers, and tiles all of them in one launch on the given stream. The C-ABI entry points
// (abi.cpp) split larger batches into several launches.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define XDN_MAX_LIST 4  // reprojected buffers per launch

namespace xdn {

// Cross-regression family (variance, regression, recon). Unused pointers may be null.
struct CrFrame {
    const float *color, *color2;  // raw colour (only for the fused log1p prologue)
    float *y1, *y2;               // log colour (input, or output of the prologue)
    float *var1, *var2;           // [H][W]
    const float *alb, *nrm;
    float *beta;                  // [H/s][W/s][2][18]
    float *outa, *outb;           // [H/rs][W/rs][3]
    int H, W;
};

hipError_t launch_variance(int n, const CrFrame *fr, int which, bool log_in, hipStream_t st);
hipError_t launch_regression(int n, const CrFrame *fr, int win, int s, hipStream_t st);
hipError_t launch_recon(int n, const CrFrame *fr, int win, int s, int rs, hipStream_t st);

// Backward of cross_regression with respect to its inputs (recon_scale 1). var1/var2/beta hold
// the recomputed forward quantities; gq, u, gcen, evar are scratch.
struct CrgFrame {
    const float *y1, *y2, *alb, *nrm;
    const float *var1, *var2, *beta;     // [H][W], [H][W], [Hs][Ws][2][18]
    const float *gouta, *goutb;          // [H][W][3]
    const float *gvara, *gvarb;          // [H][W] or null
    float4 *gq;                          // [H][W][2]
    float *u;                            // [Hs][Ws][2][18]
    float *gcen;                         // [Hs][Ws][2][12]
    float *evar;                         // [H][W][6]
    float *gya, *gyb, *galb, *gnrm;      // [H][W][3]
    int H, W;
};
hipError_t launch_cross_regression_grad(int n, const CrgFrame *fr, int win, int s, hipStream_t st);

// Spatial filter family.
struct SfFrame {
    const float *ia, *ib, *alb, *nrm, *band;          // band: [H][W][5]
    float *accA, *accB, *wsA, *wsB;                   // forward outputs
    const float *gaccA, *gaccB, *gwA, *gwB;           // upstream gradients
    float *gband;                                     // [H][W][5]
    float *gia, *gib, *galb, *gnrm;                   // input gradients
    int H, W;
};

hipError_t launch_sf_forward(int n, const SfFrame *fr, int win, hipStream_t st);
hipError_t launch_sf_grad_band(int n, const SfFrame *fr, int win, hipStream_t st);
hipError_t launch_sf_grad_inputs(int n, const SfFrame *fr, int win, hipStream_t st);

// Fused per-frame training kernel: spatial filter forward + temporal accumulation + loss +
// analytic backward to the raw network output (scripts/main.py:214-248, 291-302).
struct TrainFrame {
    const float *ols1, *ols2, *alb, *nrm;
    const float *param;                 // [H][W][6] raw network output
    const float *succ_e, *succ_prev;    // [H][W]
    const float *prev_out, *opacity;    // [H][W][3], [H][W]
    const float *prev_ols1, *prev_ols2; // screened history (temporal labels)
    float *out_log, *out;               // [H][W][3]
    float *dparam;                      // [H][W][6] d mean-loss / d param
    float *loss;                        // [H][W][3] or null
    float *loss_part;                   // one partial sum per tile
    int H, W;
};
hipError_t launch_filter_train(int n, const TrainFrame *fr, int win, hipStream_t st);
int filter_train_tiles(int H, int W);

// Reprojection.
struct RpFrame {
    const float *mvec, *linZ, *prevLinZ, *nrm, *prevNrm, *pnFw, *opac, *prevOpac;
    uint8_t *success;
    const float *in[XDN_MAX_LIST];
    float *out[XDN_MAX_LIST];
    int H, W;
};
// unit_opacity: opac and prevOpac are all ones (not read).
hipError_t launch_reproject(int n, const RpFrame *fr, int n_list, const int *dims, bool unit_opacity,
                            hipStream_t st);

// Temporal screening (optionally with the 3x3 erosion of a u8 success mask fused in).
struct TsFrame {
    const float *c1, *c2, *prev, *prev1, *prev2;
    const float *rsucc;          // eroded success (float) when not fused
    const uint8_t *succ_mask;    // raw reprojection success when erosion is fused
    float *o, *o1, *o2, *osucc;
    float *succ_e;               // eroded success output when fused (may be null)
    int H, W;
};
hipError_t launch_screening(int n, const TsFrame *fr, int win_curr, int win_prev, bool fuse_erosion,
                            hipStream_t st);

}  // namespace xdn
