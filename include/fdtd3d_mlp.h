/*
 * fdtd3d_mlp.h -- the transmission surrogate of the taper design loop (no counterpart in the
 * reference).
 *
 * A multilayer perceptron n_in -> width (x depth hidden layers, GELU) -> n_out, trained and
 * evaluated on the GPU by hand-written BF16 matrix-core (MFMA) kernels with FP32 accumulation:
 * weights, activations and back-propagated errors enter each GEMM as BF16 (round to nearest
 * even); hidden pre-activations, the output, gradients and the Adam master parameters are FP32.
 *
 * Parameter vector layout (FP32): for each layer l = 0..depth, W_l [fout][fin] row-major, then
 * b_l [fout]; fin = n_in for layer 0 and width otherwise, fout = width except n_out for the last.
 * Inputs x [n][n_in], outputs / targets y, t [n][n_out] (row-major FP32).  The loss is the mean
 * over samples and outputs of (y - t)^2.  Every function returns 0 on success and a negative
 * code on failure (message: fdtd_dev_last_error()).
 */
#ifndef FDTD3D_MLP_H
#define FDTD3D_MLP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdtd_mlp fdtd_mlp;

fdtd_mlp *fdtd_mlp_new(int n_in, int width, int depth, int n_out);
void fdtd_mlp_free(fdtd_mlp *m);
long long fdtd_mlp_param_count(const fdtd_mlp *m);
/* set (also resets the Adam moments) / read the FP32 parameters */
int fdtd_mlp_set_params(fdtd_mlp *m, const float *params);
int fdtd_mlp_get_params(fdtd_mlp *m, float *params);
/* inference: y = f(x) for n samples */
int fdtd_mlp_forward(fdtd_mlp *m, const float *x, int n, float *y);
/* loss and its gradient with respect to every parameter (grad in the parameter layout) */
int fdtd_mlp_loss_grad(fdtd_mlp *m, const float *x, const float *t, int n, double *loss, float *grad);
/* one Adam step on the batch (x, t); *loss is the loss before the step */
int fdtd_mlp_adam_step(fdtd_mlp *m, const float *x, const float *t, int n, float lr, float beta1, float beta2,
                       float eps, double *loss);
/* `steps` full-batch Adam steps on (x, t), uploaded once and kept on the device; the loss before
 * the first step and after the last one (either pointer may be NULL; each read synchronises once) */
int fdtd_mlp_train(fdtd_mlp *m, const float *x, const float *t, int n, int steps, float lr, float beta1, float beta2,
                   float eps, double *loss_first, double *loss_last);

/* Measurement.  gemm_timing(on) clears the record and, while on, brackets every matrix-core GEMM
 * launch with events recorded at its dispatch; gemm_times returns the number of recorded launches,
 * their summed duration (ms) and their arithmetic (2 x multiply-adds), and clears the record.
 * timer_start / timer_stop time everything enqueued on the network's stream in between. */
int fdtd_mlp_gemm_timing(fdtd_mlp *m, int on);
int fdtd_mlp_gemm_times(fdtd_mlp *m, int *launches, double *ms, double *flop);
int fdtd_mlp_timer_start(fdtd_mlp *m);
int fdtd_mlp_timer_stop(fdtd_mlp *m, double *elapsed_ms);

#ifdef __cplusplus
}
#endif
#endif
