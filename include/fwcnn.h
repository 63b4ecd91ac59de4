/* fwcnn.h -- C ABI of the learned coarse-grid correction (MI355X, gfx950).
 *
 * No counterpart in the reference: FUNWAVE-TVD has no learned component.
 * A coarse run (grid spacing 2*dx) is corrected after every time step by a
 * small convolutional network that predicts the one-step discrepancy
 * (d_eta, d_u, d_v) between the coarse solver and the 2x2-averaged fine
 * solver, from the coarse state (eta, u, v, depth, mask).
 *
 * Network (fixed): four 3x3 convolutions with zero padding,
 *   5 -> 32 (ReLU) -> 32 (ReLU) -> 32 (ReLU) -> 3,
 * bf16 operands on the matrix cores (v_mfma_f32_16x16x32_bf16) with fp32
 * accumulation, bf16 activations between layers, fp32 master parameters.
 *
 * Parameter vector (fp32, FW_CNN_NPARAMS values), in this order:
 *   W1[9][5][32], b1[32], W2[9][32][32], b2[32], W3[9][32][32], b3[32],
 *   W4[9][32][3], b4[3]
 * with the tap index t = 3*(dy+1) + (dx+1) for the input offset (dy, dx),
 * i.e. out[y][x][co] = b[co] + sum_t sum_ci in[y+dy][x+dx][ci] * W[t][ci][co].
 *
 * Tensors crossing the boundary are plain fp32 arrays in host memory,
 * planar per sample: inputs [B][5][H][W], outputs / targets [B][3][H][W].
 * All functions return 0 or a negative FW_E_* code (fwgpu.h).
 */
#ifndef FWCNN_H
#define FWCNN_H

#include "fwgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { FW_CNN_NPARAMS = 9 * 5 * 32 + 32 + 2 * (9 * 32 * 32 + 32) + 9 * 32 * 3 + 3 };

typedef struct fw_cnn fw_cnn;

/* Buffers for batches of up to max_batch samples of H x W cells. */
int fw_cnn_create(int H, int W, int max_batch, fw_cnn **out);
int fw_cnn_destroy(fw_cnn *net);
const char *fw_cnn_strerror(const fw_cnn *net);
int fw_cnn_set_params(fw_cnn *net, const float *params, int n);
int fw_cnn_get_params(fw_cnn *net, float *params, int n);

/* out = net(in) for B samples (host arrays). */
int fw_cnn_forward(fw_cnn *net, const float *in, float *out, int B);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the mean squared
 * error over the B samples' H*W*3 outputs against target; *loss receives
 * that mean before the update.  grad (optional, FW_CNN_NPARAMS floats)
 * receives the gradient of the loss. */
int fw_cnn_train_step(fw_cnn *net, const float *in, const float *target, int B, float lr, float *loss,
                      float *grad);

/* Device milliseconds per forward pass over B resident samples (the
 * current input buffer), averaged over iters passes timed with HIP events
 * on the network's stream. */
int fw_cnn_time_forward(fw_cnn *net, int B, int iters, double *ms);

/* In-place correction of a coarse solver state whose interior is H x W:
 * features (eta*fs[0], u*fs[1], v*fs[2], depth*fs[3], mask) of the current
 * state, network output o, then eta += o0*os[0], u += o1*os[1],
 * v += o2*os[2] on wet interior cells, with H, HU, HV, Ubar, Vbar updated
 * to match (Ubar += H*du, Vbar += H*dv).  The work is queued on the solver's
 * stream. */
int fw_cnn_correct(fw_cnn *net, fw_ctx *solver, const float fs[4], const float os[3]);

/* With on != 0, fw_cnn_correct takes the wet-cell mean of the eta
 * correction out of it (accumulated in FP64), so that the correction moves
 * no water volume; u and v are corrected as they are.  Default off. */
int fw_cnn_set_conserve(fw_cnn *net, int on);

#ifdef __cplusplus
}
#endif
#endif
