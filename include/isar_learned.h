/*
 * isar_learned.h — learned warm start for the fast-entropy autofocus (plain C ABI).
 *
 * No counterpart in the reference. The reference's autofocus (phase_adjustment.cu:282-324)
 * starts from the Doppler-tracking phase and runs 120 fixed-point iterations. Here a small
 * multilayer perceptron predicts, from the phase change of the first iteration, the remaining
 * change to the converged phase; the iteration then continues from that prediction
 * (isar_run_device with warm_start and stage_mode BACK). The network runs on the matrix cores
 * with BF16 weights and FP32 accumulation; training uses the backward and Adam entry points
 * below. All pointers named d_* are device pointers; calls taking a stream are asynchronous.
 * Every function returns 0 on success and a negative isar_status code on failure
 * (isar_last_error() has the reason).
 */
#ifndef ISAR_LEARNED_H_
#define ISAR_LEARNED_H_

#include <stddef.h>
#include <stdint.h>

#include "isar_imaging_ext.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct isar_mlp isar_mlp;

/* A network of n_layers dense layers: widths[0] inputs, widths[n_layers] outputs, ReLU after
 * every layer but the last. Buffers for batches of up to max_batch rows. */
ISAR_API int isar_mlp_create(int n_layers, const int* widths, int max_batch, isar_mlp** out);
ISAR_API int isar_mlp_destroy(isar_mlp* net);
/* Number of FP32 parameters. Layout: for each layer, W (in x out, row-major) then b (out). */
ISAR_API int isar_mlp_param_count(const isar_mlp* net);
/* Host FP32 parameters in and out; set also refreshes the BF16 weight copies and clears the Adam
 * moments. Synchronous. */
ISAR_API int isar_mlp_set_params(isar_mlp* net, const float* h_params);
ISAR_API int isar_mlp_get_params(const isar_mlp* net, float* h_params);
/* Gradients of the last backward pass, same layout as the parameters. Synchronous. */
ISAR_API int isar_mlp_get_grads(const isar_mlp* net, float* h_grads);
/* d_out[batch x widths[n]] = net(d_in[batch x widths[0]]). The network keeps d_in and its layer
 * activations for the next backward call. */
ISAR_API int isar_mlp_forward(isar_mlp* net, int batch, const float* d_in, float* d_out, void* stream);
/* Parameter gradients for the loss gradient d_grad_out[batch x widths[n]] of the last forward. */
ISAR_API int isar_mlp_backward(isar_mlp* net, int batch, const float* d_grad_out, void* stream);
/* One Adam step (bias-corrected, step >= 1) on the FP32 master parameters. */
ISAR_API int isar_mlp_adam_step(isar_mlp* net, float lr, float beta1, float beta2, float eps, int step,
                                void* stream);
/* Mean squared error of d_pred against d_target (rows x cols) per row into d_loss_rows (double,
 * each row's share of the mean) and, when d_grad is not NULL, its gradient. */
ISAR_API int isar_mse(int rows, int cols, const float* d_pred, const float* d_target, float* d_grad,
                      double* d_loss_rows, void* stream);

/* Per-pulse phase change d_out[f out_stride + 2m] + j d_out[f out_stride + 2m + 1] =
 * phi_b[f][m] conj(phi_a[f][m]) for frames x M complex phases (the network's inputs and target);
 * out_stride >= 2M floats per frame. */
ISAR_API int isar_phase_change(int frames, int M, const void* d_phi_a, const void* d_phi_b, float* d_out,
                               int out_stride, void* stream);
/* Warm start phi_w[f][m] = phi_b[f][m] * unit(pred[f][2m] + j pred[f][2m+1]). */
ISAR_API int isar_compose_phase(int frames, int M, const void* d_phi_b, const float* d_pred, void* d_phi_w,
                                void* stream);
/* Entropy -sum p ln p, p = img^2 / sum img^2, of `frames` images of n floats each (double). */
ISAR_API int isar_image_entropy(int frames, int n, const float* d_img, double* d_out, void* stream);
/* The plain autofocus judges the network: per frame, d_out[f] = d_phi_b[f] when its image
 * entropy d_ent_b[f] is lower than d_ent_a[f], else d_phi_a[f] (frames x M complex phases), and
 * d_ent_out[f] (optional) the entropy of the phase kept. With phi_a the plain phase after the
 * front iterations and phi_b a warm start, a prediction that does not lower the entropy is
 * dropped and the plain iteration continues. Chained, it keeps the lowest of several candidates
 * (the first on ties). */
ISAR_API int isar_select_phase(int frames, int M, const void* d_phi_a, const double* d_ent_a, const void* d_phi_b,
                               const double* d_ent_b, void* d_out, double* d_ent_out, void* stream);
/* The final judgement (the plain autofocus as fallback and judge, phase_adjustment.cu:302-324):
 * per frame, d_out[f] = the learned image d_img_learned[f] when its entropy d_ent_learned[f] is
 * no worse than the plain result's d_ent_plain[f] (within rel_tol relative), else the plain
 * image d_img_plain[f]; d_accepted[f] (optional) = 1 where the learned image was kept. Images
 * are frames x n floats; d_out may alias either input. */
ISAR_API int isar_accept_image(int frames, int n, const float* d_img_plain, const double* d_ent_plain,
                               const float* d_img_learned, const double* d_ent_learned, double rel_tol, float* d_out,
                               int32_t* d_accepted, void* stream);
/* Half step from phi_a toward phi_b per pulse: phi_a sqrt(phi_b conj(phi_a)) on the shorter arc
 * (a second warm-start candidate, half of the predicted change). */
ISAR_API int isar_half_phase(int frames, int M, const void* d_phi_a, const void* d_phi_b, void* d_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ISAR_LEARNED_H_ */
