/* C ABI of the warm-start MLP (MI355X, gfx950).
 *
 * A 3-layer perceptron n_in -> hidden -> hidden -> n_out with GELU
 * activations, trained with Adam on mean-squared error.  Every matrix
 * product runs on MFMA with BF16 operands and FP32 accumulation; master
 * weights, activations and optimiser state are FP32.  It proposes starting
 * points for the batched interior-point engine (include/qtqp_batch.h,
 * qtqp_batch_set_start); it has no counterpart in the reference, whose only
 * starting point is the default one of src/qtqp/__init__.py:299-307.
 *
 * Parameters are one packed FP32 vector: W1 (n_in x hidden, row-major), b1,
 * W2 (hidden x hidden), b2, W3 (hidden x n_out), b3.  Host arrays are
 * caller-owned and copied during each call; calls are synchronous.
 */
#ifndef QTQP_MLP_H_
#define QTQP_MLP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qtqp_mlp qtqp_mlp;

int qtqp_mlp_create(int32_t n_in, int32_t hidden, int32_t n_out, int32_t max_batch,
                    qtqp_mlp** out);
int64_t qtqp_mlp_num_params(qtqp_mlp* h);
int qtqp_mlp_set_params(qtqp_mlp* h, const float* params);
int qtqp_mlp_get_params(qtqp_mlp* h, float* params);

/* y (batch x n_out) = network(x (batch x n_in)); any batch size. */
int qtqp_mlp_forward(qtqp_mlp* h, const float* x, int64_t batch, float* y);

/* Starting points of the batched interior-point engine from the network, on
 * the device.  x: batch x n_in standardised inputs (host); tm, ts: the
 * n_out = n + 2 m + 2 target means and scales (host, FP64).  Each row's
 * outputs are un-standardised in FP64 (raw = out ts + tm) and read as
 * (x, log y, log s, log tau, log kappa): x_start (batch x n), y_start and
 * s_start (batch x m: exp of the clipped logs on the m - z inequality rows,
 * y = raw and s = 0 on the z equality rows) and tau_start (batch) are
 * written to DEVICE memory, packed as qtqp_batch_set_start_device takes
 * them.  A row with a non-finite output gets the default start of
 * src/qtqp/__init__.py:299-307 and valid[row] = 0 (valid is host memory
 * and may be NULL).  No reference counterpart (the reference has only the
 * default start). */
int qtqp_mlp_propose(qtqp_mlp* h, const float* x, int64_t batch, const double* tm,
                     const double* ts, int32_t n, int32_t m, int32_t z, double* x_start,
                     double* y_start, double* s_start, double* tau_start, int32_t* valid);
/* Mean-squared-error loss of one batch (batch <= max_batch) and its gradient
 * with respect to the packed parameters (no update). */
int qtqp_mlp_loss_grad(qtqp_mlp* h, const float* x, const float* t, int32_t batch, float* grad,
                       double* loss);

/* Uploads the training set (count rows of x and t) once and runs `steps`
 * Adam steps on mini-batches drawn without replacement per epoch (shuffled
 * with `seed`); loss_per_step (steps values) may be NULL. */
int qtqp_mlp_train(qtqp_mlp* h, const float* x, const float* t, int64_t count, int32_t batch,
                   int32_t steps, float lr, float beta1, float beta2, float eps, uint64_t seed,
                   float* loss_per_step);

void qtqp_mlp_free(qtqp_mlp* h);
const char* qtqp_mlp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QTQP_MLP_H_ */
