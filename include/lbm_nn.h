/*
 * lbm_nn.h -- C ABI of the learned warm start for the 2D cavity solver.
 *
 * No reference counterpart: the reference (zzhao0930/MPI-Multi-Process-Version)
 * always starts from rest with a linear temperature profile
 * (natural-convection_2dim_mono.f90:199-240).  This network predicts a
 * steady (u, v, T) field from the cavity's parameters so that
 * lbm_init_fields() can start the solver close to its steady state; the
 * solver then runs to the same stopping rule, so it remains the judge of the
 * result.
 *
 * Network: a convolutional encoder-decoder on a G x G node grid (G a
 * multiple of 8) with base width W (16 unless created with
 * lbm_nn_create_width).  Input channels: log10(Ra), Pr, and the node
 * coordinates x/H and y/H (the stretched grid), each mapped affinely onto
 * [-1, 1] over its trained range by the caller; output channels: u/V, v/V
 * and T, with V = sqrt(g beta H) / 4 a quarter of the cavity's free-fall
 * velocity (LB units).
 *   encoder: 3x3 conv 4->W, 3x3 stride-2 convs W->2W->4W->4W (ReLU)
 *   decoder: three stages of x2 nearest upsampling, concatenation with the
 *            encoder feature map of the same size, 3x3 conv (ReLU):
 *            (4W+4W)->4W, (4W+2W)->2W, (2W+W)->W
 *   head:    1x1 conv W->3 (linear)
 * That is 657 W^2 + 57 W + 3 parameters: 169 107 at W = 16, 516 687 at
 * W = 28.
 * Every convolution is an im2col product on the matrix cores: BF16 weights
 * and activations, FP32 accumulation; master weights, biases, gradients and
 * the Adam state are FP32.
 *
 * Host arrays are FP32, sample-major: x[b][4][G][G], y[b][3][G][G] with
 * the grid's i index fastest.  Parameters are listed layer by layer, each
 * layer's weights [out][in][ky][kx] then its biases.
 */
#ifndef LBM_NN_H
#define LBM_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_NN_IN_CHANNELS  4
#define LBM_NN_OUT_CHANNELS 3

typedef struct lbm_nn_ctx *lbm_nn;

/* grid: G (a multiple of 8, at least 16); batch_max: the largest batch a
 * call may pass; seed: weight initialisation (He-uniform, zero biases). */
int32_t lbm_nn_create(int32_t grid, int32_t batch_max, uint64_t seed, int32_t device,
                      lbm_nn *out);
/* The same with base width `width` (a multiple of 4 from 4 to 64). */
int32_t lbm_nn_create_width(int32_t grid, int32_t width, int32_t batch_max, uint64_t seed,
                            int32_t device, lbm_nn *out);
int32_t lbm_nn_destroy(lbm_nn net);
int32_t lbm_nn_param_count(lbm_nn net, int64_t *count);
int32_t lbm_nn_get_params(lbm_nn net, float *params);
/* Sets the FP32 master weights and clears the Adam state. */
int32_t lbm_nn_set_params(lbm_nn net, const float *params);

/* Inference: y = net(x) for `batch` samples. */
int32_t lbm_nn_forward(lbm_nn net, int32_t batch, const float *x, float *y);

/* Mean squared error over all outputs and its gradient with respect to the
 * parameters (param_count values), without changing the parameters. */
int32_t lbm_nn_loss_grad(lbm_nn net, int32_t batch, const float *x, const float *target,
                         float *loss, float *grad);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the mean squared
 * error of the batch; loss receives the loss before the step. */
int32_t lbm_nn_train_step(lbm_nn net, int32_t batch, const float *x, const float *target,
                          float lr, float *loss);

#ifdef __cplusplus
}
#endif

#endif /* LBM_NN_H */
