/*
 * xcorr_cnn.h -- C ABI of the learned offset proposal network (MI355X, gfx950).
 *
 * No counterpart in the reference (diego-romano/gmtsar-gpu has no learned
 * component).  A small CNN looks at a 4 x 4 box-downsampled amplitude pair of
 * a 256 x 256 xcorr window (-xsearch 64 -ysearch 64) and predicts the window's
 * offset and a confidence; the proposal pipeline uses it to run a shrunken
 * search (xsearch 16, centred on the prediction) where it is confident and the
 * plain search elsewhere.  Plain xcorr produces its training labels and judges
 * its error.
 *
 * Network (all 3 x 3 convolutions, zero padding, ReLU, 32 output channels,
 * BF16 weights and activations, FP32 accumulation on the matrix cores):
 *   input  x0[64][64][2]    normalised master / slave crops, bf16
 *   conv1  2 -> 32, stride 2     -> 32 x 32
 *   conv2  32 -> 32, stride 2    -> 16 x 16
 *   conv3  32 -> 32, stride 2    ->  8 x  8
 *   conv4  32 -> 32, stride 1    ->  8 x  8
 *   head   global mean over 8 x 8, FP32 linear 32 -> 3:
 *          o[0], o[1] = crop offset (x, y) in downsampled pixels, o[2] = confidence logit
 * Parameters: one FP32 vector of XCORR_CNN_NPARAM values, per conv W[9][32][cin]
 * (tap = 3 ky + kx, then output channel, then input channel) followed by b[32],
 * then Wh[3][32], bh[3].
 *
 * Every entry point returns 0 or a negative XCORR_HIP_E* code (xcorr_hip.h);
 * xcorr_cnn_last_error() gives the message.  Device pointers are the caller's.
 */
#ifndef XCORR_CNN_H
#define XCORR_CNN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XCORR_CNN_NPARAM 28451

typedef struct xcorr_cnn xcorr_cnn;

/* Context for up to max_batch training samples per step (inference takes any n). */
int xcorr_cnn_create(int device, int max_batch, xcorr_cnn **out);
void xcorr_cnn_destroy(xcorr_cnn *net);
const char *xcorr_cnn_last_error(const xcorr_cnn *net);
void *xcorr_cnn_stream(xcorr_cnn *net);
/* Queue all later work on `stream` (a hipStream_t; NULL is the null stream,
 * the default stream of most callers), so that it is ordered with the caller's
 * other work on that stream.  Until the first call the context's own stream
 * is used. */
int xcorr_cnn_set_stream(xcorr_cnn *net, void *stream);
int xcorr_cnn_sync(xcorr_cnn *net);

/* Host FP32 parameters in / out.  set_params also clears the optimiser state. */
int xcorr_cnn_set_params(xcorr_cnn *net, const float *params);
int xcorr_cnn_get_params(xcorr_cnn *net, float *params);

/* 4 x 4 box mean of the amplitude of an image (device, ny lines of nx samples,
 * XCORR_HIP_CINT16 or XCORR_HIP_FLOAT32) into out (device, (ny/4) x (nx/4) floats). */
int xcorr_cnn_box4(xcorr_cnn *net, const void *img, int nx, int ny, int format, float *out);

/* Normalised 64 x 64 crops x0[n][64][64][2] (device bf16) of the downsampled
 * master / slave images at corners[n] = {mx, my, sx, sy} (device int32, in
 * downsampled pixels; outside the images reads as zero).  Each channel is
 * shifted to zero mean and scaled to unit variance (scale 1e6 when flat). */
int xcorr_cnn_crops(xcorr_cnn *net, const float *ds_m, int dnx_m, int dny_m, const float *ds_s, int dnx_s,
                    int dny_s, const int32_t *corners, int n, uint16_t *x0);

/* Forward pass of n inputs into out[n][3] (device): o[0], o[1], o[2].
 * infer: one fused kernel per input (activations in LDS); forward: the
 * layer-by-layer training kernels (same arithmetic). */
int xcorr_cnn_infer(xcorr_cnn *net, const uint16_t *x0, int n, float *out);
int xcorr_cnn_forward(xcorr_cnn *net, const uint16_t *x0, int n, float *out);

/* Crops and fused inference in one kernel, straight from the downsampled images. */
int xcorr_cnn_propose(xcorr_cnn *net, const float *ds_m, int dnx_m, int dny_m, const float *ds_s, int dnx_s,
                      int dny_s, const int32_t *corners, int n, float *out);

/* Training targets per sample (device float[n][5]): tx, ty (crop offset in
 * downsampled pixels), tc (confidence label 0/1), wr (regression weight), wc
 * (confidence weight).  Loss = mean over the batch of
 *   wr (huber(o0 - tx) + huber(o1 - ty)) + wc bce(sigmoid(o2), tc).
 * gradient: loss and its gradient (host float[XCORR_CNN_NPARAM]) without an
 * update; out (device float[n][3]) may be NULL.
 * train_step: one Adam step (beta 0.9 / 0.999, eps 1e-8) with learning rate
 * lr; *loss (host, may be NULL: no wait) receives the batch loss. */
int xcorr_cnn_gradient(xcorr_cnn *net, const uint16_t *x0, const float *tgt, int n, float *loss, float *grad,
                       float *out);
int xcorr_cnn_train_step(xcorr_cnn *net, const uint16_t *x0, const float *tgt, int n, float lr, float *loss);

#ifdef __cplusplus
}
#endif
#endif
