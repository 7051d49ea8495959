/*
 * xdn.h -- C ABI of libxdn.so, the gfx950 (MI355X) implementation of cross-denoiser's per-frame
 * filtering path.
 *
 * Conventions
 *   - Every image is a contiguous NHWC float32 device buffer [B, H, W, C] (C stated per argument),
 *     the layout of the reference ops (scripts/ops.py). B images are processed independently.
 *   - `stream` is a hipStream_t (NULL = default stream). All work is enqueued on it; no entry
 *     point synchronizes, allocates device memory or keeps device state between calls (the one
 *     exception, xdn_filter_train_noalloc, uses a fixed array of the library instead of a
 *     workspace), so every call can be captured in a hipGraph.
 *   - Return value: 0 on success, XDN_EINVAL for bad arguments (nothing launched), XDN_EHIP for a
 *     HIP launch error. xdn_last_error() gives the message of the last failure on the calling
 *     thread; argument messages repeat the reference's text where it has one.
 *   - Pointers may not alias unless stated.
 *   - An empty batch (batch 0, with valid H, W and attributes) returns 0 without reading or
 *     writing any buffer, so its pointers may be NULL (the data pointers of zero-element tensors);
 *     xdn_cnn5_backward and xdn_unet_backward, whose gradient has no batch dimension, write zeros
 *     to grad.
 *   - Frame size: the cnn5 entry points and cnn5 slots take frames of fewer than 2^25 pixels (the
 *     network kernels address their 64 B/px activations with 32-bit offsets; 7680x4320 fits), the
 *     spatial-filter entry points and every slot fewer than 2^32; larger ones are XDN_EINVAL.
 */
#ifndef XDN_H
#define XDN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XDN_OK 0
#define XDN_EINVAL (-1)
#define XDN_EHIP (-2)

/* Library version string. */
const char *xdn_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char *xdn_last_error(void);

/* ---------------------------------------------------------------------------------------------
 * CrossRegression op: replaces REGISTER_OP("CrossRegression") / CrossRegressionOp::Compute
 * (ops/regression.cc:25-57, 125-196) and CrossRegressionFunc (ops/regression.cu.cc:477-500).
 *   imga, imgb, albedo, normal  [B,H,W,3]
 *   outa, outb                  [B,H/recon_scale,W/recon_scale,3]
 *   out_vara, out_varb          [B,H,W,1]
 *   workspace                   >= xdn_cross_regression_workspace_size(B,H,W,sparsity) bytes
 *                               (the regression coefficients; replaces the static betaA/betaB
 *                               temporaries of regression.cc:174-180)
 * Errors (InvalidArgument in the reference): recon_scale > sparsity (regression.cc:134-135),
 * H or W not divisible by recon_scale (:159-160); additionally H, W divisible by sparsity and
 * H, W > win_size/2 (the reference reads unwritten or out-of-bounds memory otherwise).
 * ------------------------------------------------------------------------------------------- */
size_t xdn_cross_regression_workspace_size(int batch, int height, int width, int sparsity);
int xdn_cross_regression(const float *imga, const float *imgb, const float *albedo, const float *normal,
                         float *outa, float *outb, float *out_vara, float *out_varb, void *workspace,
                         size_t workspace_bytes, int batch, int height, int width, int win_size, int sparsity,
                         int recon_scale, void *stream);

/* Same op with prepare_input (scripts/main.py:207-212) fused in: takes the raw colours, writes
 * y1 = log1p(color), y2 = log1p(color2) [B,H,W,3] and then runs the op on them. */
int xdn_cross_regression_from_color(const float *color, const float *color2, const float *albedo,
                                    const float *normal, float *y1, float *y2, float *outa, float *outb,
                                    float *out_vara, float *out_varb, void *workspace, size_t workspace_bytes,
                                    int batch, int height, int width, int win_size, int sparsity, void *stream);

/* Backward of the CrossRegression op with respect to imga, imgb, albedo and normal, for
 * recon_scale 1. The reference registers no gradient for this op (ops/regression.cc:25-57 and
 * scripts/ops.py:23-32 register one only for SpatialFilter); this is the adjoint of
 * CrossRegressionFunc (ops/regression.cu.cc:244-500) for torch.autograd. Every window weight is
 * recomputed; no forward state is passed in.
 *   grad_outa, grad_outb [B,H,W,3];  grad_vara, grad_varb [B,H,W,1] (either may be NULL = zero)
 *   grad_imga, grad_imgb, grad_albedo, grad_normal [B,H,W,3] (overwritten)
 * The workspace is caller-owned, xdn_cross_regression_grad_workspace_size() bytes. */
size_t xdn_cross_regression_grad_workspace_size(int batch, int height, int width, int sparsity);
int xdn_cross_regression_grad(const float *imga, const float *imgb, const float *albedo, const float *normal,
                              const float *grad_outa, const float *grad_outb, const float *grad_vara,
                              const float *grad_varb, float *grad_imga, float *grad_imgb, float *grad_albedo,
                              float *grad_normal, void *workspace, size_t workspace_bytes, int batch, int height,
                              int width, int win_size, int sparsity, void *stream);

/* ---------------------------------------------------------------------------------------------
 * SpatialFilter op: replaces REGISTER_OP("SpatialFilter") / SpatialFilterOp::Compute
 * (ops/regression.cc:59-86, 235-282) and SpatialFilterFunc (ops/regression.cu.cc:502-517).
 *   imga, imgb, albedo, normal  [B,H,W,3];  band [B,H,W,5]
 *   out_acca, out_accb          [B,H,W,3];  out_wgtsuma, out_wgtsumb [B,H,W,1]
 * No per-tap state is stored (the reference writes 2x[B,H,W,16] gradBand temporaries here).
 * ------------------------------------------------------------------------------------------- */
int xdn_spatial_filter(const float *imga, const float *imgb, const float *albedo, const float *normal,
                       const float *band, float *out_acca, float *out_accb, float *out_wgtsuma,
                       float *out_wgtsumb, int batch, int height, int width, int win_size, void *stream);

/* SpatialFilterGrad op: replaces REGISTER_OP("SpatialFilterGrad") / SpatialFilterGradOp::Compute
 * (ops/regression.cc:88-103, 283-323) and SpatialFilterGradFunc (ops/regression.cu.cc:519-534).
 * grad_band [B,H,W,5] = d L / d band, recomputing the window weights instead of reading moments
 * stored by the forward pass. */
int xdn_spatial_filter_grad(const float *grad_acca, const float *grad_accb, const float *grad_wgtsuma,
                            const float *grad_wgtsumb, const float *imga, const float *imgb, const float *img_albedo,
                            const float *img_normal, const float *band, float *out_grad_band, int batch, int height,
                            int width, int win_size, void *stream);

/* Gradient of SpatialFilter w.r.t. its image inputs, which the reference's registered gradient
 * leaves as None (scripts/ops.py:23-32). Any of the four outputs [B,H,W,3] may be NULL. */
int xdn_spatial_filter_grad_inputs(const float *grad_acca, const float *grad_accb, const float *grad_wgtsuma,
                                   const float *grad_wgtsumb, const float *imga, const float *imgb,
                                   const float *img_albedo, const float *img_normal, const float *band,
                                   float *grad_imga, float *grad_imgb, float *grad_albedo, float *grad_normal,
                                   int batch, int height, int width, int win_size, void *stream);

/* ---------------------------------------------------------------------------------------------
 * Reproject op: replaces REGISTER_OP("Reproject") / ReprojectOp::Compute (ops/reproject.cc:34-69,
 * 87-178) and ReprojectFunc (ops/reproject.cu.cc:240-273).
 *   mvec, linear_z, prev_linear_z, normal, prev_normal, pn_fwidth [B,H,W,3]
 *   opacity, prev_opacity [B,H,W,1];  success [B,H,W,1] bool (1 byte)
 *   inputs[l] / outputs[l] [B,H,W,dims[l]] for l < n_list, 1 <= dims[l] <= 4; the pointer
 *   arrays are host memory read during the call (no device pointer table, no host sync).
 * ------------------------------------------------------------------------------------------- */
int xdn_reproject(const float *mvec, int n_list, const float *const *inputs, const int *dims,
                  const float *linear_z, const float *prev_linear_z, const float *normal, const float *prev_normal,
                  const float *pn_fwidth, const float *opacity, const float *prev_opacity, uint8_t *success,
                  float *const *outputs, int batch, int height, int width, void *stream);

/* ---------------------------------------------------------------------------------------------
 * TemporalScreening op: replaces REGISTER_OP("TemporalScreening") / TemporalScreeningOp::Compute
 * (ops/temporal.cu.cc:204-285).
 *   current1, current2, prev, prev1, prev2 [B,H,W,3]; reproj_success [B,H,W,1] float
 *   output, output1, output2 [B,H,W,3]; success [B,H,W,1] float
 * ------------------------------------------------------------------------------------------- */
int xdn_temporal_screening(const float *current1, const float *current2, const float *prev, const float *prev1,
                           const float *prev2, const float *reproj_success, float *output, float *output1,
                           float *output2, float *success, int batch, int height, int width, int win_curr,
                           int win_prev, void *stream);

/* The same op preceded by the 3x3 erosion of the boolean reprojection mask
 * (scripts/main.py:270-276, tf.nn.erosion2d + 1): takes `success_mask` [B,H,W,1] u8 and also
 * writes the eroded mask to `eroded` [B,H,W,1] float (may be NULL). */
int xdn_erode_temporal_screening(const float *current1, const float *current2, const float *prev,
                                 const float *prev1, const float *prev2, const uint8_t *success_mask, float *eroded,
                                 float *output, float *output1, float *output2, float *success, int batch,
                                 int height, int width, int win_curr, int win_prev, void *stream);

/* ---------------------------------------------------------------------------------------------
 * Fused spatiotemporal filtering step of train_step (scripts/main.py:291-302): the SpatialFilter
 * forward with band = 1/(x^2+1e-4), alpha = sigmoid(x) of the raw network output `param`
 * [B,H,W,6] (main.py:110-113), aggregation_loss (main.py:214-248) and its gradient back to
 * `param`. Replaces SpatialFilter + aggregation_loss + SpatialFilterGrad + the band/alpha
 * head's backward. Labels follow main.py:283-286 (label1 = ols2, label2 = ols1,
 * prev_label1 = prev_ols2, prev_label2 = prev_ols1).
 *   ols1, ols2, albedo, normal, prev_out, prev_ols1, prev_ols2 [B,H,W,3]
 *   success (eroded), success_prev, opacity [B,H,W,1]
 *   out_log, out [B,H,W,3]; dparam [B,H,W,6] = d mean(loss of the image) / d param;
 *   loss [B,H,W,3] (may be NULL); avg_loss [B] (mean loss per image; may be NULL, then no loss
 *   sum is launched)
 *   workspace >= xdn_filter_train_workspace_size(B,H,W) bytes
 * ------------------------------------------------------------------------------------------- */
size_t xdn_filter_train_workspace_size(int batch, int height, int width);
int xdn_filter_train(const float *ols1, const float *ols2, const float *albedo, const float *normal,
                     const float *param, const float *success, const float *success_prev, const float *prev_out,
                     const float *opacity, const float *prev_ols1, const float *prev_ols2, float *out_log, float *out,
                     float *dparam, float *loss, float *avg_loss, void *workspace, size_t workspace_bytes, int batch,
                     int height, int width, int win_size, void *stream);

/* The same op with no workspace argument and no device allocation. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The per-tile loss partials go to a fixed
 * array inside the library, one per GPU, held until the process ends:
 * XDN_FILTER_TRAIN_SCRATCH_BYTES = 1 MiB, i.e. batch x ceil(H/16) x ceil(W/16) <= 262,144 tiles
 * (e.g. batch 16 at 1920x1080, or one 8192x8192 image). Larger inputs return XDN_EINVAL before
 * anything is launched. Calls on one GPU share that array: calls that overlap on the device
 * (other streams, other threads) race and must be serialized by the caller. */
#define XDN_FILTER_TRAIN_SCRATCH_BYTES (1u << 20)
int xdn_filter_train_noalloc(const float *ols1, const float *ols2, const float *albedo, const float *normal,
                             const float *param, const float *success, const float *success_prev,
                             const float *prev_out, const float *opacity, const float *prev_ols1,
                             const float *prev_ols2, float *out_log, float *out, float *dparam, float *loss,
                             float *avg_loss, int batch, int height, int width, int win_size, void *stream);

/* Keras Adam step (tf.keras.optimizers.Adam, scripts/main.py:81,303) on a flat parameter buffer:
 * m += (g-m)(1-b1); v += (g^2-v)(1-b2); p -= alpha m/(sqrt(v)+eps) with
 * alpha = lr sqrt(1-b2^step)/(1-b1^step), step counted from 1. */
int xdn_adam(float *param, const float *grad, float *m, float *v, int count, int step, float lr, float beta1,
             float beta2, float eps, void *stream);

/* Frame report of the reference's driver loop (scripts/main.py:344-346, RelL2 of :191-196), in
 * one pass: out_composed = out + emissive + env_light (the displayed frame; may be NULL), the
 * reference ref + ref_emissive + ref_env_light, and err[b] = mean over pixels and channels of
 * (ref_c - out_c)^2 / (mean over channels of ref_c)^2 + 1e-2). All images [B,H,W,3]; emissive,
 * env_light, ref_emissive, ref_env_light may be NULL (zero). The workspace holds per-block
 * partial sums, xdn_frame_error_workspace_size(B,H,W) bytes. */
size_t xdn_frame_error_workspace_size(int batch, int height, int width);
int xdn_frame_error(const float *out, const float *emissive, const float *env_light, const float *ref,
                    const float *ref_emissive, const float *ref_env_light, float *out_composed, float *err,
                    void *workspace, size_t workspace_bytes, int batch, int height, int width, void *stream);

/* ---------------------------------------------------------------------------------------------
 * cnn5: the bandwidth network of the benchmark configuration (the reference's online network is
 * the Keras U-Net of scripts/main.py:83-118, trained per frame at :289-303): five 3x3 'same'
 * convolutions 15 -> 32 -> 32 -> 32 -> 32 -> 6, ReLU between them, BF16 MFMA operands, FP32
 * accumulation. The network input is the concatenation [ols1, ols2, prev_out, albedo, normal]
 * (main.py:291), fused into the first layer; its raw 6-channel output feeds xdn_filter_train.
 *   params   FP32 [xdn_cnn5_param_count()]: per layer W [3][3][ci][co] (Keras HWIO) then b [co]
 *   packed   xdn_cnn5_packed_bytes() bytes of bf16 MFMA fragments, rebuilt by xdn_cnn5_pack after
 *            every parameter update
 *   workspace xdn_cnn5_workspace_size(B,H,W) bytes: the four hidden activations kept for the
 *            backward pass, two gradient buffers and the weight-gradient partial sums
 * A batch of B images shares one set of weights; xdn_cnn5_backward writes the gradient summed over
 * the batch into grad [param_count] (dparam = d loss / d network output, [B,H,W,6]).
 * ------------------------------------------------------------------------------------------- */
size_t xdn_cnn5_param_count(void);
size_t xdn_cnn5_packed_bytes(void);
size_t xdn_cnn5_workspace_size(int batch, int height, int width);
int xdn_cnn5_pack(const float *params, void *packed, void *stream);
int xdn_cnn5_forward(const float *ols1, const float *ols2, const float *prev_out, const float *albedo,
                     const float *normal, const void *packed, float *param, void *workspace, size_t workspace_bytes,
                     int batch, int height, int width, void *stream);
int xdn_cnn5_backward(const float *ols1, const float *ols2, const float *prev_out, const float *albedo,
                      const float *normal, const float *dparam, const void *packed, float *grad, void *workspace,
                      size_t workspace_bytes, int batch, int height, int width, void *stream);

/* ---------------------------------------------------------------------------------------------
 * unet: the reference's own bandwidth network (scripts/main.py:83-118 unet_simple + the 1x1 head
 * of :106-108; 31,414 parameters): conv3x3 15->8->8, 2x2 max pool, 8->16->16, pool, 16->32->32,
 * transposed 1x1 stride-2 conv 32->32, concat with the 16-channel skip, conv3x3 48->16->16,
 * transposed conv 16->16, concat with the 8-channel skip, conv3x3 24->8, conv1x1 8->6 (FP32).
 * ReLU after every 3x3 convolution; the transposed convolutions are linear. 3x3 layers run on
 * BF16 MFMA with FP32 accumulation, the 1x1 layers in FP32; activations are stored as bf16.
 * Same calling convention as cnn5 (the two are interchangeable in front of xdn_filter_train):
 *   params   FP32 [xdn_unet_param_count()] in Keras order, per layer kernel then bias: 3x3
 *            kernels [3][3][ci][co], transposed kernels [co][ci] (Keras 1x1xCOxCI), head [ci][co]
 *   packed   xdn_unet_packed_bytes(), rebuilt by xdn_unet_pack after every parameter update
 *   workspace xdn_unet_workspace_size(B,H,W) bytes; height and width must be multiples of 4
 * xdn_unet_workspace_layout gives the offset and size of workspace buffer `index` for
 * inspection: 0-12 the activations c1_1, c1_2, pool1, c2_1, c2_2, pool2, c3_1, c3_2, c3_3, c4_1,
 * c4_2, c4_3, c5_1 (bf16 [B][h][w][ch]); 13-25 the backward gradients in the order they are made.
 * ------------------------------------------------------------------------------------------- */
size_t xdn_unet_param_count(void);
size_t xdn_unet_packed_bytes(void);
size_t xdn_unet_workspace_size(int batch, int height, int width);
int xdn_unet_workspace_layout(int batch, int height, int width, int index, size_t *offset, size_t *bytes);
int xdn_unet_pack(const float *params, void *packed, void *stream);
int xdn_unet_forward(const float *ols1, const float *ols2, const float *prev_out, const float *albedo,
                     const float *normal, const void *packed, float *param, void *workspace, size_t workspace_bytes,
                     int batch, int height, int width, void *stream);
int xdn_unet_backward(const float *ols1, const float *ols2, const float *prev_out, const float *albedo,
                      const float *normal, const float *dparam, const void *packed, float *grad, void *workspace,
                      size_t workspace_bytes, int batch, int height, int width, void *stream);

/* ---------------------------------------------------------------------------------------------
 * Batched online denoising step: train_step (scripts/main.py:250-313) for many render streams.
 * One call advances every slot by one frame, each stage being ONE launch over all slots (their
 * resolutions may differ): cross-regression with the log1p prologue, reprojection of the history
 * (opacities one, main.py:266), erosion + temporal screening (win 7/1), network forward (cnn5 or
 * the reference's U-Net, per slot.network; one call takes slots of one network), fused filter +
 * temporal accumulation + loss + backward, network backward, Keras Adam. The cnn5 Adam update
 * rewrites the bf16 fragment slots of every weight it updates (the U-Net's fragments are rebuilt by
 * one packing launch), so the step never rebuilds the whole fragment buffer from params.
 * Labels and history handoff follow main.py:283-286 and :338-341: the caller passes this frame's
 * out_log/ols1/ols2 buffers as the next frame's prev_* (ping-pong) together with this frame's
 * linear_z/normal. Frame 0 of a stream uses zero history (main.py:323-330).
 * No counterpart call exists in the reference (it runs one stream, op by op).
 * ------------------------------------------------------------------------------------------- */
typedef struct xdn_slot {
    int height, width; /* divisible by 4, > 8 */
    /* this frame, NHWC float32: [H,W,3] except opacity [H,W,1] */
    const float *color, *color2, *albedo, *normal, *mvec, *linear_z, *pn_fwidth, *opacity;
    /* history from the previous frame: out_log, ols1, ols2, linear_z, normal [H,W,3] */
    const float *prev_out_log, *prev_ols1, *prev_ols2, *prev_linear_z, *prev_normal;
    /* outputs of this frame [H,W,3] (out = max(0, expm1(out_log))) */
    float *out_log, *ols1, *ols2, *out;
    /* slot-owned scratch of xdn_slot_scratch_bytes(H, W) bytes, prepared once by xdn_slot_init */
    void *scratch;
    /* network state: params, grad, Adam m and v [param count of the slot's network], packed
     * fragments [packed bytes of the network]. `packed` must come from xdn_slot_init (or
     * xdn_cnn5_pack / xdn_unet_pack), and be packed again after any write to `params` from outside
     * the step: the step only updates the fragments of the weights it changes. */
    float *params, *grad, *adam_m, *adam_v;
    void *packed;
    int adam_step; /* step number of this update, from 1 (ignored when adam_t is set) */
    float *avg_loss; /* [1] mean loss of this frame */
    /* optional device int32 step counter (number of updates done so far). When set, each step
     * increments it and forms the Adam bias correction on the device, so a captured step graph
     * stays valid on replay; NULL uses adam_step from the host. */
    int *adam_t;
    int network; /* XDN_NET_CNN5 (xdn_cnn5_*) or XDN_NET_UNET (xdn_unet_*) */
} xdn_slot;
#define XDN_NET_CNN5 0
#define XDN_NET_UNET 1

size_t xdn_slot_scratch_bytes(int height, int width);
/* Byte offsets inside the scratch of its XDN_SLOT_NBUF intermediate buffers, in this order:
 * y1, y2 [H,W,3]; var1, var2 [H,W,1]; regression coefficients; reprojection success [H,W,1] u8;
 * reprojected out_log, ols1, ols2 [H,W,3]; screened prev_out, prev_ols1, prev_ols2 [H,W,3];
 * success_prev, eroded success [H,W,1]; dparam, network output [H,W,6]; ones [H,W,1];
 * loss partials; network workspace (sized for either network); Adam step size [1]. For
 * inspection and tests. */
#define XDN_SLOT_NBUF 20
int xdn_slot_scratch_offsets(int height, int width, size_t *offsets);
/* Fills the slot's constant scratch and packs its network; call once when the slot is admitted. */
int xdn_slot_init(const xdn_slot *slot, void *stream);
/* One frame for each of n_slots (<= 16) slots, enqueued on `stream`. */
int xdn_denoise_step(const xdn_slot *slots, int n_slots, float learning_rate, void *stream);

/* The same step in parts, so that consecutive frames of a stream can overlap on two HIP streams.
 * XDN_STAGE_REGRESS: cross-regression of this frame (reads only the frame; writes ols1/ols2 and
 * slot scratch used by nothing else). XDN_STAGE_REPROJECT: reprojection of the history
 * (prev_out_log/prev_ols1/prev_ols2). XDN_STAGE_FINISH: screening, network, filter, loss,
 * backward, Adam. FINISH needs both others done; REGRESS of frame t+1 writes the buffers frame t
 * reads as history, so it may start once frame t's REPROJECT is done. xdn_denoise_step ==
 * all three on one stream. No counterpart in the reference (its train_step is one TF call chain). */
#define XDN_STAGE_REGRESS 1
#define XDN_STAGE_REPROJECT 2
#define XDN_STAGE_FINISH 4
#define XDN_STAGE_ALL 7
int xdn_denoise_stages(const xdn_slot *slots, int n_slots, float learning_rate, int stages, void *stream);

/* The same step captured once into a HIP graph (one graph per frame: the slots' buffer
 * addresses are baked in, so a caller keeps one graph per distinct set of frame / history
 * buffers, e.g. per history parity). Every slot must have adam_t set. `stream` must not be the
 * null stream; the graph is launched with xdn_denoise_graph_launch and freed with
 * xdn_denoise_graph_destroy. No counterpart in the reference (its steps are TF eager calls,
 * scripts/main.py:250-313). */
int xdn_denoise_graph_create(const xdn_slot *slots, int n_slots, float learning_rate, void *stream, void **graph);
int xdn_denoise_graph_launch(void *graph, void *stream);
int xdn_denoise_graph_destroy(void *graph);

#ifdef __cplusplus
}
#endif

#endif /* XDN_H */
