// Launchers of the small per-frame training kernels (internal header).
#pragma once
#include <hip/hip_runtime.h>

namespace xdn {

// One flat parameter tensor updated by Adam.
struct AdamJob {
    float *param, *grad, *m, *v;
    int count;
    float alpha;              // lr * sqrt(1 - b2^t) / (1 - b1^t)
    const float *alpha_dev;   // when set, the step size is read from here instead
};

hipError_t launch_adam(int n, const AdamJob *jobs, float b1, float b2, float eps, hipStream_t st);
// Per job: t = ++(*counter[i]); alpha[i] = lr * sqrt(1 - b2^t) / (1 - b1^t) (double, rounded once).
hipError_t launch_adam_tick(int n, int *const *counter, float *const *alpha, float lr, float b1, float b2,
                            hipStream_t st);
hipError_t launch_sums(int n, const float *const *parts, const int *count, const float *scale, float *const *out,
                       hipStream_t st);

}  // namespace xdn
