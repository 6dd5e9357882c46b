/*
 * opensps_emu.h -- the C ABI of the learned emulator in libopensps_hip.so: an MLP (n_in -> layers x hidden GELU ->
 * n_out) that maps a model's parameters to the PCA coefficients of its log-spectrum and its band magnitudes,
 * trained on spectra the GPU batch itself generates. Forward, backward and Adam are gfx950 kernels on the BF16
 * matrix cores (FP32 accumulation, FP32 master weights and Adam moments).
 *
 * No reference counterpart: the reference's only network on the GPU is cue's nebular-emission MLP
 * (src/accel/hip/opensps_hip_kernels.hip:3670-3841, src/physics/opensps_cue.f90), an inference-only layer pair;
 * this emulator is the learned fast path in front of opensps_hip_csp_batch, which stays its fallback and its judge.
 *
 * Conventions: arrays are row-major per sample (the features of a sample contiguous: x[i n_in + k]); pointers
 * documented "device" are HIP device addresses, "host" ones host memory. Statuses as opensps_hip.h (0 success,
 * 1 bad handle, 3 bad argument, 10 a HIP runtime error; opensps_hip_last_error(NULL, ...) gives the message).
 * Parameters ("the flat vector") are FP32 in layer order: W_0 (hidden x n_in, row-major), b_0 (hidden), W_1 ...,
 * W_L (n_out x hidden), b_L.
 */
#ifndef OPENSPS_EMU_H
#define OPENSPS_EMU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OpsEmu opensps_emu_t;

/* n_in <= 64, n_out <= 256, hidden a multiple of 128 (<= 1024), layers >= 1 hidden layers; weights drawn with
 * seed (He-uniform), biases 0, scalings the identity. */
int opensps_emu_create(int device, int n_in, int n_out, int hidden, int layers, unsigned seed, opensps_emu_t **emu);
int opensps_emu_destroy(opensps_emu_t *emu);
/* The number of parameters of the flat vector. */
int opensps_emu_nparams(opensps_emu_t *emu, long *n);
/* Host FP32 flat vectors. set also clears the Adam moments and step count. */
int opensps_emu_set_params(opensps_emu_t *emu, const float *params);
int opensps_emu_get_params(opensps_emu_t *emu, float *params);
/* Host FP32: inputs x' = (x - in_mu) / in_sd; outputs y = y' out_sd + out_mu (y' the network's). */
int opensps_emu_set_scaling(opensps_emu_t *emu, const float *in_mu, const float *in_sd, const float *out_mu,
                            const float *out_sd);
/* Per-output weights of the training loss (host FP32, n_out; default 1). */
int opensps_emu_set_loss_weights(opensps_emu_t *emu, const float *w);
/* Adam's constants (defaults 0.9, 0.999, 1e-8). */
int opensps_emu_set_adam(opensps_emu_t *emu, float beta1, float beta2, float eps);

/* Inference: device x (n x n_in) -> device y (n x n_out), asynchronous on stream (a hipStream_t; NULL: the
 * emulator's stream). */
int opensps_emu_predict(opensps_emu_t *emu, int n, const float *x, float *y, void *stream);
/* The training loss L = sum over samples and outputs of w_o (y'_o - t'_o)^2 / n (t' the scaled targets) and its
 * gradient with respect to the flat vector, without an update: device x (n x n_in), t (n x n_out); host grads
 * (nparams) and loss, or NULL. Synchronous. */
int opensps_emu_gradients(opensps_emu_t *emu, int n, const float *x, const float *t, float *grads, float *loss);
/* One Adam step on a minibatch (device x, t as above) with learning rate lr; *loss (host, or NULL) receives the
 * minibatch loss before the step. Asynchronous unless loss is given. */
int opensps_emu_train_step(opensps_emu_t *emu, int n, const float *x, const float *t, float lr, float *loss);
/* Runs the emulator's work on stream (a hipStream_t; NULL: the null stream) instead of its own, so that it is
 * ordered with the caller's work on its inputs and outputs. */
int opensps_emu_set_stream(opensps_emu_t *emu, void *stream);
/* Waits for the emulator's stream. */
int opensps_emu_sync(opensps_emu_t *emu);

#ifdef __cplusplus
}
#endif
#endif
