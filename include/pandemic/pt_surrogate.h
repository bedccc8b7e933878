/*
 * pt_surrogate.h — C ABI of the epidemic surrogate: an ensemble of MLPs that
 * maps a 16-dimensional scenario vector to 2 strains x 120 days x 3
 * log-incidence statistics (mean, 10% and 90% quantiles over replicates).
 *
 * No counterpart in the reference (it has no surrogate); this is the
 * scenario-screening model trained on the replicate-batch engine's output
 * (pandemicthrust_amd/ensemble.py).  Architecture per member:
 * 16 -> 512 -> 512 -> 512 -> 512 (GELU after each) -> 720, BF16 weights and
 * activations, FP32 accumulation (gfx950 MFMA), FP32 master weights and Adam
 * moments, mean-squared-error loss.  All members train on the same batches
 * in the same kernel launches.
 *
 * Parameter layout per member (floats): for each layer l = 0..4 the weight
 * matrix W_l (in_l rows x out_l columns, row-major) then the bias b_l.
 */
#ifndef PANDEMIC_PT_SURROGATE_H
#define PANDEMIC_PT_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_MLP_IN 16
#define PT_MLP_HIDDEN 512
#define PT_MLP_LAYERS 5
#define PT_MLP_OUT 720

typedef struct pt_mlp pt_mlp;

/* Floats per member (weights and biases of the 5 layers). */
int64_t pt_mlp_param_count(void);

/* members >= 1, batch = rows per training step (multiple of 64). */
int pt_mlp_create(int32_t members, int32_t batch, int32_t device, pt_mlp** out);

/* Master parameters of every member (members * pt_mlp_param_count floats);
 * setting them also resets the Adam state. */
int pt_mlp_set_params(pt_mlp* m, const float* params);
int pt_mlp_get_params(pt_mlp* m, float* params);

/* Training set resident in HBM: X[n][16] (inputs), Y[n][720] (targets). */
int pt_mlp_set_data(pt_mlp* m, const float* X, const float* Y, int64_t n);

/* steps Adam steps (beta1 0.9, beta2 0.999, eps 1e-8, learning rate lr);
 * step s trains every member on rows order[s*batch .. (s+1)*batch) of the
 * training set.  loss_out[s*members + k] (may be NULL) = member k's batch
 * loss before step s. */
int pt_mlp_train(pt_mlp* m, const int32_t* order, int32_t steps, float lr, float* loss_out);

/* Loss and gradients of every member on one batch of training rows (no
 * update): grads_out[members * pt_mlp_param_count], loss_out[members]. */
int pt_mlp_gradients(pt_mlp* m, const int32_t* rows, float* grads_out, float* loss_out);

/* Predictions of every member for X[n][16]: out[(k*n + i)*720 + j]. */
int pt_mlp_predict(pt_mlp* m, const float* X, int64_t n, float* out);

/* Training steps are replayed from a HIP graph of one step captured per
 * pt_mlp_train call (default), or enqueued kernel by kernel (on = 0); the
 * results are identical. */
int pt_mlp_use_graphs(pt_mlp* m, int32_t on);

/* Milliseconds of the last pt_mlp_train (HIP events around the steps). */
double pt_mlp_last_train_ms(pt_mlp* m);

void pt_mlp_destroy(pt_mlp* m);

#ifdef __cplusplus
}
#endif

#endif /* PANDEMIC_PT_SURROGATE_H */
