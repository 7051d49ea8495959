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
};

hipError_t launch_adam(int n, const AdamJob *jobs, float b1, float b2, float eps, hipStream_t st);
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

}  // namespace xdn
