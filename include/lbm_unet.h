/*
 * lbm_unet.h -- C ABI of the learned warm-start: a 3-D convolutional U-Net that
 * predicts the developed (rho, u) field from a solid mask and a Reynolds
 * number, trained and run on gfx950 by this library's own BF16 MFMA
 * convolution kernels (forward, input gradient, weight gradient).
 *
 * The reference has no counterpart: it always starts from a uniform
 * equilibrium (initialize_f, src/Initialization.jl:57-83).  The network's
 * output only feeds lbm_init_equilibrium(); the solver's own drag residual
 * still decides convergence, so the network can shorten a run but not change
 * its converged answer.
 *
 * Architecture: levels l = 0..L-1 with base * 2^l channels; two 3x3x3 convs
 * (ReLU) per level, 2x2x2 average pooling down, 2x2x2 stride-2 transposed
 * convs (ReLU) up, skip connections by channel concatenation, 1x1x1 head.
 * Activations BF16 (channels last), FP32 accumulation, FP32 master weights,
 * Adam.  Host tensors are FP32, layout [batch][channel][z][y][x].
 */
#ifndef LBM_UNET_H
#define LBM_UNET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t nx, ny, nz;  /* spatial size; each a multiple of 2^(levels-1) */
    int32_t batch;       /* largest batch; batch * nx * ny * nz < 2^31 and that times base / 8 < 2^32 */
    int32_t cin;         /* input channels (<= 8), e.g. 2: solid mask, Re */
    int32_t cout;        /* output channels (<= 16), e.g. 4: rho, ux, uy, uz */
    int32_t base;        /* channels at level 0 (multiple of 16), e.g. 16 */
    int32_t levels;      /* 2..5 */
    uint64_t seed;       /* weight initialisation */
    float lr, beta1, beta2, eps;
    int32_t periodic;    /* bit 0 x, bit 1 y, bit 2 z: the 3x3x3 convolutions wrap around along these axes
                            instead of padding with zeros (a periodic domain); needs every level at least
                            16 x 4 x 4 voxels (x, y, z), else lbm_unet_create fails */
} lbm_unet_config;

typedef struct lbm_unet lbm_unet;

int lbm_unet_create(const lbm_unet_config* cfg, lbm_unet** out);
int lbm_unet_destroy(lbm_unet* net);
int lbm_unet_last_error(const lbm_unet* net, char* buf, size_t n);
/* Number of FP32 master parameters. */
int lbm_unet_num_params(lbm_unet* net, int64_t* n);
/* Master parameters in the library's layout (see lbm_unet_param_layout). */
int lbm_unet_get_params(lbm_unet* net, float* out);
int lbm_unet_set_params(lbm_unet* net, const float* in);
/* Layer table: for layer i, kind (0 conv3, 1 up2, 2 head), cin, cout, offset of
 * its weights [npar][cout][taps][cin] and of its bias [cout] in the flat
 * parameter vector.  Returns the number of layers in *nlayers. */
int lbm_unet_param_layout(lbm_unet* net, int32_t* nlayers, int32_t* kind, int32_t* cin, int32_t* cout,
                          int64_t* w_off, int64_t* b_off);
/* Forward pass: x [batch][cin][N] -> y [batch][cout][N]. */
int lbm_unet_forward(lbm_unet* net, const float* x, int32_t batch, float* y);
/* One Adam step on the mean-squared error; returns the loss (before the step). */
int lbm_unet_train_step(lbm_unet* net, const float* x, const float* target, int32_t batch, float* loss);
/* Gradient of the loss w.r.t. the master parameters at the current weights
 * (no update), for gradient checks. */
int lbm_unet_gradient(lbm_unet* net, const float* x, const float* target, int32_t batch, float* loss,
                      float* grad);
/* Adam learning rate for the following train steps (a schedule's knob). */
int lbm_unet_set_lr(lbm_unet* net, float lr);
/* Timing of the last train step / forward (ms, device events) and its FLOP count. */
int lbm_unet_last_timing(lbm_unet* net, double* ms, double* flops);

#ifdef __cplusplus
}
#endif

#endif /* LBM_UNET_H */
