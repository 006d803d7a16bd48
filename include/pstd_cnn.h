/* Warm-start network of the MI355X PSTD core: C ABI.
 *
 * None of these entry points has a counterpart in the reference, which has
 * no learned component (its driver always starts from the incident field,
 * PSTD_QScat.cpp:959-1118).  The network predicts, from the potential and
 * the incidence, the difference between the steady-state field amplitude
 * and the initial field, so that a solve can start near its steady state.
 *
 * Architecture: a 3D convolutional encoder-decoder on an R^3 grid (R a
 * multiple of 8): 3x3x3 convolutions (stride 1, zero padding 1) with leaky
 * ReLU (slope 0.1), three 2x average-pooling levels with 16, 32, 64 and 128
 * channels, 2x nearest upsampling with additive skips, 1,770,320
 * parameters.  Weights and activations are bfloat16 on MFMA with FP32
 * accumulation; master weights, biases, gradients and the Adam state are
 * FP32.  Tensors are channels-last: [batch][z][y][x][channel].
 *
 * Parameter order (get/set): for each of the 16 convolutions in network
 * order, W[27][Cout][Cin] (tap t = 9 (dz+1) + 3 (dy+1) + (dx+1)) then
 * b[Cout].
 *
 * Errors as for the solver (pstd.h): 22 bad argument, 12 out of memory,
 * 19 no gfx950 device; pstd_last_error() gives the message.
 */
#ifndef PSTD_CNN_H
#define PSTD_CNN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSTD_CNN_IN_CHANNELS 16  /* input channels per voxel (padded) */
#define PSTD_CNN_OUT_CHANNELS 2  /* predicted channels (Re, Im) */

typedef struct pstd_cnn pstd_cnn;

/* Network for batches of `batch` samples on a resolution^3 grid, weights
 * initialised with He-normal draws of a SplitMix64 stream from `seed`,
 * zero biases, fresh Adam state. */
int pstd_cnn_create(int resolution, int batch, unsigned long long seed, pstd_cnn **out);
void pstd_cnn_destroy(pstd_cnn *net);

/* Number of parameters (1,770,320). */
long pstd_cnn_num_params(void);

/* FP32 master parameters, in the order above.  set also resets the Adam
 * moments and step count. */
int pstd_cnn_get_params(pstd_cnn *net, float *out, size_t n);
int pstd_cnn_set_params(pstd_cnn *net, const float *in, size_t n);

/* Forward pass.  input: [batch][R][R][R][16] floats (rounded to bfloat16
 * on the device); output: [batch][R][R][R][2] floats. */
int pstd_cnn_forward(pstd_cnn *net, const float *input, float *output);

/* Forward and backward pass with the loss mean((out[..., 0:2] - target)^2);
 * target [batch][R][R][R][2].  grad (may be NULL) receives the FP32
 * gradient of the loss in parameter order; loss (may be NULL) the loss.
 * Parameters are not changed. */
int pstd_cnn_gradients(pstd_cnn *net, const float *input, const float *target, float *grad, float *loss);

/* Forward, backward and one Adam step (beta1 0.9, beta2 0.999, eps 1e-8;
 * Kingma and Ba, Algorithm 1) with learning rate lr. */
int pstd_cnn_train_step(pstd_cnn *net, const float *input, const float *target, float lr, float *loss);

#ifdef __cplusplus
}
#endif

#endif
