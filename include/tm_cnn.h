/*
 * C ABI of the proposal network: a small convolutional network, trained on the exhaustive search's
 * own peaks, that proposes particle locations and coarse orientation bins for a local exhaustive
 * refinement (tm_refine_score_batch in tm_refine.h). It has no counterpart in the reference, which
 * always runs the exhaustive search.
 *
 * Network: n_layers 3x3 convolutions (padding 1, stride 1 or 2, bias, ReLU) over a square
 * single-channel image of input_size (standardised per image on the device), then one 1x1 head
 * with 1 + n_bins outputs per cell of the last feature map: a location logit (column 0) and
 * n_bins orientation logits. Activations are FP32 and channels-last; every product runs on bf16
 * operands with FP32 accumulation (MFMA 16x16x16); master weights and Adam moments are FP32.
 *
 * Parameters, flattened in this order (get/set): for each conv layer l, weights [C_l][3][3][C_{l-1}]
 * (k = (ky * 3 + kx) * C_{l-1} + c, the input channel fastest) and bias [C_l]; then the head's
 * weights [1 + n_bins][C_L] and bias [1 + n_bins].
 *
 * Every call returns TM_OK (0) or a negative TM_ERR_* code from tm_hip.h; tm_cnn_last_error()
 * describes the last failure on the calling thread.
 */
#ifndef TM_CNN_H
#define TM_CNN_H

#include "tm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TM_CNN_MAX_LAYERS 16

typedef struct tm_cnn tm_cnn;

typedef struct tm_cnn_config {
  int input_size;                     /* S: the image is S x S (S divisible by the total stride)     */
  int n_layers;                       /* 3x3 convolutions, 1 .. TM_CNN_MAX_LAYERS                     */
  int channels[TM_CNN_MAX_LAYERS];    /* output channels of each convolution                          */
  int strides[TM_CNN_MAX_LAYERS];     /* 1 or 2                                                       */
  int n_bins;                         /* orientation bins of the head                                 */
  int max_batch;                      /* images per forward / backward call                           */
} tm_cnn_config;

int tm_cnn_create(int device_ordinal, const tm_cnn_config* config, tm_cnn** out);
int tm_cnn_destroy(tm_cnn* h);

/* Number of parameters and the edge G of the output grid (cells of S / G pixels). */
int tm_cnn_shape(tm_cnn* h, long long* n_params, int* grid);

/* Parameters as one flat FP32 array (the order above). init_params draws them on the host from a
 * seeded generator (He-normal weights, zero biases); set/get copy them in and out. Adam moments
 * and the step counter are reset by init and set. */
int tm_cnn_init_params(tm_cnn* h, unsigned long long seed);
int tm_cnn_set_params(tm_cnn* h, const float* params);
int tm_cnn_get_params(tm_cnn* h, float* params);

/* Forward pass over batch images (batch x S x S, host, row-major). logits (may be NULL) receives
 * batch x G x G x (1 + n_bins) values, cell-major. */
int tm_cnn_forward(tm_cnn* h, const float* images, int batch, float* logits);

/* The same forward pass over batch images already in device memory (batch x S x S FP32 on the
 * handle's device, not modified), enqueued on the handle's stream with no copy and no wait: for
 * training loops whose inputs stay resident on the GPU. */
int tm_cnn_forward_device(tm_cnn* h, const float* images, int batch);

/* Location logits of the last forward pass (batch x G x G), and the 1 + n_bins logits of n chosen
 * cells (flat indices b G^2 + y G + x) into out (n x (1 + n_bins)): what a proposal step reads
 * instead of the whole logit volume. */
int tm_cnn_heat(tm_cnn* h, float* heat);
int tm_cnn_cell_logits(tm_cnn* h, const int* cells, int n, float* out);

/* Loss and gradients of the last forward pass. heat_target: batch x G x G values in {0, 1};
 * bin_target: batch x G x G bins (-1: no orientation term at that cell). Loss = mean over cells of
 * the location logit's binary cross-entropy (positives weighted by pos_weight) + the mean over
 * cells with a bin of the orientation softmax cross-entropy. loss[2] receives both terms; grads
 * (may be NULL) the flat gradient. */
int tm_cnn_backward(tm_cnn* h, const float* heat_target, const int* bin_target, float pos_weight, float* loss,
                    float* grads);

/* The same backward pass from targets in device memory (heat_target FP32, bin_target int32, batch x
 * G x G), enqueued with no wait. n_labelled: the number of cells with a bin >= 0 (the orientation
 * term is their mean); bins >= n_bins count as none here (tm_cnn_backward rejects them).
 * tm_cnn_loss returns the two loss terms of the last backward pass, waiting for it;
 * tm_cnn_synchronize waits for all enqueued work of the handle. */
int tm_cnn_backward_device(tm_cnn* h, const float* heat_target, const int* bin_target, long long n_labelled,
                           float pos_weight);
int tm_cnn_loss(tm_cnn* h, float* loss);
int tm_cnn_synchronize(tm_cnn* h);

/* One Adam step with the gradients of the last backward pass (bias-corrected moments), enqueued on
 * the handle's stream: every later call on the handle sees the updated parameters. */
int tm_cnn_adam_step(tm_cnn* h, float lr, float beta1, float beta2, float eps);

/* Test hook: the post-ReLU output of conv layer l (0-based; l = n_layers gives the head's logits)
 * of the last forward pass, batch x S_l x S_l x C_l, channels-last. */
int tm_cnn_get_activation(tm_cnn* h, int layer, float* out);

/* Test hook: C[M][N] = A[M][K] . B[N][K]^T on bf16-rounded operands (A, B FP32 host arrays,
 * rounded to bf16 on the device, K any), FP32 accumulation. */
int tm_cnn_debug_gemm(int device_ordinal, const float* A, const float* B, int M, int N, int K, float* C);

const char* tm_cnn_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TM_CNN_H */
