// Internal interface of the warm-start CNN kernels (pdflow_cnn.hip) for the host layer
// (pdflow_cnn_host.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pdf {
namespace cnn {

// 3x3 same convolution: x [nb][h][w][cin] bf16, wt [coutp][kp] bf16, bias [cout] f32 or null;
// output channel co < cout goes to y16 (bf16) and/or y32 (f32) at row p, column yoff + co
// with row stride ystride
struct ConvArgs {
    const uint16_t *x, *wt;
    const float *bias;
    uint16_t *y16;
    float *y32;
    int nb, h, w, cin, cout, coutp, kp, ystride, yoff, relu;
};

// weight gradient: dwt [coutp][kp] f32 += sum_p dz[p][co] * im2col(x)[p][k]
struct WGradArgs {
    const uint16_t *x, *dz;
    float *dwt;
    int nb, h, w, cin, coutp, kp;
};

hipError_t probe(const uint16_t *A, const uint16_t *B, float *C, hipStream_t s);
hipError_t conv3x3(const ConvArgs &a, hipStream_t s);          // LDS-staged (used)
hipError_t conv3x3_direct(const ConvArgs &a, hipStream_t s);   // one 64x32 tile per workgroup, no LDS
hipError_t conv3x3_wgrad(const WGradArgs &a, hipStream_t s);
hipError_t bias_grad(const uint16_t *dz, float *db, int P, int coutp, int cout, hipStream_t s);
hipError_t relu_grad(const uint16_t *dy, const uint16_t *y, uint16_t *dz, size_t n, hipStream_t s);
hipError_t pool(const uint16_t *x, uint16_t *y, int nb, int h, int w, int c, hipStream_t s);
hipError_t pool_grad(const uint16_t *dy, uint16_t *dx, int nb, int h, int w, int c, bool acc, hipStream_t s);
hipError_t upsample(const uint16_t *x, uint16_t *y, int nb, int h, int w, int c, int ys, int off, hipStream_t s);
hipError_t upsample_grad(const uint16_t *dy, uint16_t *dx, int nb, int h, int w, int c, int ys, int off,
                         hipStream_t s);
hipError_t slice(const uint16_t *src, int ss, int soff, uint16_t *dst, int ds, int doff, size_t npix, int c, bool acc,
                 hipStream_t s);
hipError_t f32_to_bf16(const float *x, uint16_t *y, size_t n, hipStream_t s);
hipError_t prepare_weights(const float *wm, uint16_t *wt, uint16_t *wd, int cin, int cout, int coutp, int kp, int kdp,
                           hipStream_t s);
hipError_t l1_loss(const float *out, const float *target, const float *mask, uint16_t *grad, float *partial, int P,
                   float scale, hipStream_t s);
// dst row r = src row idx[r]; row_floats a multiple of 4, src/dst 16-byte aligned
hipError_t gather_rows(const float *src, const unsigned *idx, float *dst, unsigned count, size_t row_floats,
                       hipStream_t s);
hipError_t adam(float *w, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                float bc1, float bc2, float gscale, hipStream_t s);

}  // namespace cnn
}  // namespace pdf
