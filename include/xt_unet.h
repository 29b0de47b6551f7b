/* xt_unet.h -- learned warm start for the ICD: a 3D U-Net that maps a filtered
 * backprojection to a predicted converged MBIR volume, with BF16 MFMA forward,
 * backward and Adam kernels (FP32 accumulation, FP32 master weights).
 *
 * No counterpart in the reference: TIMBIR starts ICD from zero or from a
 * lower-resolution ICD (initICD, Source_Code_4D_Fast/XT_ICD_update.c:318-536).
 * The network only chooses the starting volume; the MAP cost and the ICD that
 * polishes the prediction are unchanged.
 *
 * Volumes are float32 [d0][d1][d2] (d2 fastest) in device memory; each of d0,
 * d1, d2 must be a multiple of 8 (three 2x2x2 poolings).  The ICD's own object
 * layout [N_y][N_x][N_z] is such a volume, so a warm start runs in place.
 *
 * Network (4 levels, channels C/2C/4C/8C with C = 16): two 3x3x3 convs with
 * bias and ReLU per level, 2x2x2 average pooling down, nearest-neighbour 2x
 * upsampling and skip concatenation up, a 1x1x1 head to one channel added to
 * the scaled input (out = in*scale + head(...), result / scale).
 *
 * Parameters are one float32 vector: for each of the 14 3x3x3 convs in order
 * (stem, enc1b, enc2a, enc2b, enc3a, enc3b, bott_a, bott_b, dec3a, dec3b,
 * dec2a, dec2b, dec1a, dec1b) the weights [C_out][27][C_in] (tap index
 * (a0+1)*9 + (a1+1)*3 + (a2+1) for offsets a in {-1,0,1} along d0,d1,d2) then
 * the bias [C_out]; then the head weights [C] and bias [1].
 *
 * Thread safety: one xt_unet per host thread at a time; calls are synchronous. */
#ifndef XT_UNET_H
#define XT_UNET_H

#include <stdint.h>

#include "xt_icd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xt_unet xt_unet;

/* Network for volumes of shape d0 x d1 x d2 with activations and gradients
 * allocated for that shape.  Weights He-initialised from `seed` (head zero, so
 * the initial prediction is the input).  `scale` multiplies the input volume
 * (attenuation per voxel ~1e-3 -> ~1).  `train` != 0 also allocates the
 * gradient and Adam buffers. */
int xt_unet_create(int32_t d0, int32_t d1, int32_t d2, float scale, uint64_t seed, int32_t train,
                   xt_unet **net);
void xt_unet_destroy(xt_unet *net);
int64_t xt_unet_param_count(const xt_unet *net);
/* Host copies of the float32 parameter vector; set resets the Adam moments. */
int xt_unet_get_params(xt_unet *net, float *host, int64_t n);
int xt_unet_set_params(xt_unet *net, const float *host, int64_t n);
int xt_unet_get_scale(const xt_unet *net, float *scale);
/* out = net(in); both device float32 volumes of the create shape (may alias). */
int xt_unet_forward(xt_unet *net, const float *in_dev, float *out_dev);
/* Forward + backward of loss = mean((out*scale - target*scale)^2): *loss gets
 * the loss, the gradient vector stays on the device (host copy when grad_host
 * is not null).  Requires train != 0. */
int xt_unet_loss_grad(xt_unet *net, const float *in_dev, const float *target_dev, double *loss,
                      float *grad_host);
/* Adam on the gradient of the last xt_unet_loss_grad (bias-corrected). */
int xt_unet_adam_step(xt_unet *net, float lr, float beta1, float beta2, float eps);
/* One training step: loss_grad then adam_step. */
int xt_unet_train_step(xt_unet *net, const float *in_dev, const float *target_dev, float lr,
                       double *loss);
/* The same three calls on host float32 volumes (copied through device staging
 * buffers the network allocates on first use). */
int xt_unet_forward_host(xt_unet *net, const float *in_host, float *out_host);
int xt_unet_loss_grad_host(xt_unet *net, const float *in_host, const float *target_host,
                           double *loss, float *grad_host);
int xt_unet_train_step_host(xt_unet *net, const float *in_host, const float *target_host,
                            float lr, double *loss);
/* Milliseconds of the last forward and backward (device time, HIP events). */
int xt_unet_last_ms(const xt_unet *net, double *fwd_ms, double *bwd_ms);

/* Warm start of an ICD state: filtered backprojection (xt_icd_fbp), the
 * network applied to the volume in place on the state's stream, clamp at zero
 * when the positivity constraint is on, FP64 resync of e.  The network's shape
 * must be (N_y, N_x, N_z) of a single-frame state. */
int xt_icd_unet_start(xt_icd_state *state, xt_unet *net);

#ifdef __cplusplus
}
#endif

#endif
