// Launchers of the small per-frame training kernels (internal header).
#pragma once
#include <hip/hip_runtime.h>
#include "xdn_common.h"

namespace xdn {

// One flat parameter tensor updated by Adam.
struct AdamJob {
    float *param, *grad, *m, *v;
    int count;
    float alpha;              // lr * sqrt(1 - b2^t) / (1 - b1^t)
    const float *alpha_dev;   // when set, the step size is read from here instead
    void *packed;             // when set, the updated weights also go to these bf16 fragments (layout below)
};

// Fragment layout of a network whose flat parameters are, per layer, a [9][ci][co] kernel then a
// [co] bias: per layer a forward B-fragment block frag[t][kc][lane][8] (element 8h + j of lane
// 32h + n holds W[t][16 kc + 8h + j][n]), a data-gradient block (W[8 - t][n][16 kc + 8h + j], none
// when dg_off < 0) and an FP32 bias block; offsets in bytes from the packed base. Entries outside
// ci x co are not written (they stay as the full packing left them: zero).
struct AdamPackLayer {
    int w_off, b_off, ci, co;
    int fwd_off, dg_off, bias_off;
};
constexpr int ADAM_PACK_LAYERS = 5;

hipError_t launch_adam(int n, const AdamJob *jobs, float b1, float b2, float eps, hipStream_t st,
                       const AdamPackLayer *layers = nullptr);
// Device Adam step counters: per entry t = ++(*counter[i]); alpha[i] = lr * sqrt(1 - b2^t) / (1 - b1^t)
// (double, rounded once).
struct AdamTick {
    int n;
    int *counter[XDN_MAX_FRAMES];
    float *alpha[XDN_MAX_FRAMES];
    float lr, b1, b2;
};
// out[f] = scale[f] * sum(parts[f][0 .. count[f])); with `tick` the same launch also advances those
// Adam step counters (the training step's loss sum precedes its Adam update on the stream).
hipError_t launch_sums(int n, const float *const *parts, const int *count, const float *scale, float *const *out,
                       hipStream_t st, const AdamTick *tick = nullptr);

// Frame report (displayed frame and RelL2 partial sums) of one image; see frame_error_kernel.
struct ErrFrame {
    const float *out, *emissive, *env_light;          // [H][W][3]; the last two may be null
    const float *ref, *ref_emissive, *ref_env_light;  // [H][W][3]; the last two may be null
    float *composed;                                  // [H][W][3] or null
    float *part;                                      // [frame_error_blocks(H, W)]
    int H, W;
};
int frame_error_blocks(int H, int W);
hipError_t launch_frame_error(int n, const ErrFrame *fr, hipStream_t st);
// p[0 .. count) = 0 with a kernel (no memset node when the caller captures a graph)
hipError_t launch_zero(float *p, size_t count, hipStream_t st);

}  // namespace xdn
