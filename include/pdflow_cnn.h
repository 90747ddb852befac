/* C-ABI of the learned warm start's network (part of libpdflow_hip.so).  No counterpart in the
 * reference: PD-Flow always starts from the upsampled coarser level (scene_flow_impair.cpp:124-125).
 *
 * A small BF16 U-Net (three scales, 3x3 convolutions, ~0.47 M parameters) that maps the four
 * planes I1, Z1, I2, Z2 of the level a solve starts from (NHWC, f32, normalised by the caller:
 * I/255 - 0.5, Z/4) to du, dv (pixels at that level) and 10*dw (metres), which initialise the
 * solve through pdf_solve_warm / pdf_batch_set_start_level.  Products run on MFMA with FP32
 * accumulation; parameters and the Adam state are FP32.
 *
 * Flat parameter order (pdf_net_get_params / set / grads): for each layer in the order
 * e1a, e1b, e2a, e2b, b1, b2, d2a, d2b, d1a, d1b, head: weights [cout][tap][cin]
 * (tap = 3*(dy+1)+(dx+1)), then bias [cout].
 */
#ifndef PDFLOW_CNN_H
#define PDFLOW_CNN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdf_net pdf_net;

/* buffers for batches of up to max_batch images of h x w (multiples of 4); He-normal weights
 * from seed, zero biases */
int pdf_net_create(pdf_net **out, unsigned max_batch, unsigned h, unsigned w, uint64_t seed);
int pdf_net_destroy(pdf_net *n);
const char *pdf_net_last_error(const pdf_net *n);      /* NULL n: last create failure */
int pdf_net_num_params(const pdf_net *n, size_t *count);
int pdf_net_get_params(pdf_net *n, float *host);
int pdf_net_set_params(pdf_net *n, const float *host);  /* also resets the Adam state */
int pdf_net_get_grads(pdf_net *n, float *host);        /* gradients of the last backward */
/* forward pass: x_dev [batch][h][w][4] f32 (device); y_dev [batch][h][w][3] f32 (device, may
 * be NULL: pdf_net_output_device then gives the network's own output buffer) */
int pdf_net_forward(pdf_net *n, const float *x_dev, unsigned batch, float *y_dev);
int pdf_net_output_device(pdf_net *n, const float **out);
/* after pdf_net_forward: L1 loss sum(mask * |y - target|) / (3 * pixels) and its gradients */
int pdf_net_loss_backward(pdf_net *n, const float *target_dev, const float *mask_dev, unsigned batch,
                          float *loss);
/* after pdf_net_forward: gradients for an explicit output gradient dy_dev [batch][h][w][3] */
int pdf_net_backward_from(pdf_net *n, const float *dy_dev, unsigned batch);
int pdf_net_adam_step(pdf_net *n, float lr);
/* forward + loss/backward + Adam step */
int pdf_net_train_step(pdf_net *n, const float *x_dev, const float *target_dev, const float *mask_dev,
                       unsigned batch, float lr, float *loss);
/* the same on samples idx_host[0..batch) of a device-resident data set of `total` samples laid
 * out as x_all [total][h][w][4], t_all [total][h][w][3], m_all [total][h][w] */
int pdf_net_train_step_indexed(pdf_net *n, const float *x_all, const float *t_all, const float *m_all,
                               size_t total, const unsigned *idx_host, unsigned batch, float lr, float *loss);
int pdf_net_synchronize(pdf_net *n);
/* run on the caller's HIP stream (hipStream_t) from now on; the network's own stream is
 * released and the caller's is never destroyed by the library */
int pdf_net_set_stream(pdf_net *n, void *stream);
/* weights file: 8-byte magic "PDFWNET1", uint64 count, int32 h, w, step, then the flat
 * parameters (f32, little endian) */
int pdf_net_save(pdf_net *n, const char *path);
int pdf_net_load(pdf_net *n, const char *path);
/* one v_mfma_f32_16x16x16_bf16 product C = A B of two row-major 16x16 BF16 host matrices,
 * through the operand layout the convolution kernels assume (a layout check) */
int pdf_net_mfma_probe(const uint16_t *A, const uint16_t *B, float *C);

#ifdef __cplusplus
}
#endif
#endif
