// Internal interface of the learned warm start and the per-pair quality reductions
// (pdflow_learn.hip) for the batched solver (pdflow_batch.hip).  No counterpart in the
// reference.
#pragma once

#include <hip/hip_runtime.h>

#include "pdflow_internal.hpp"

struct pdf_net;

namespace pdf {

// one pair's view of the network-size level (rows x cols, column-major planes)
struct LearnArgs {
    const float *I1, *Z1, *I2, *Z2;   // pyramid planes of both frames at that level
    unsigned rows, cols;
    float *x;                         // network input slot [rows][cols][4]
    const float *y;                   // network output slot [rows][cols][3]
    float *warm;                      // du_prev, dv_prev, dw planes (rows*cols each)
    // training labels from the finest level of a solved pair
    const float *du, *dv, *dw;        // finest total flow (after the median filter)
    unsigned frows, fcols;
    float *t, *m;                     // label slot [rows][cols][3], mask slot [rows][cols]
};

// one pair's finest level for the energy / residual reduction
struct MetricArgs {
    const float *dct, *dcu, *dcv, *ddt, *ddu, *ddv, *mu_uv, *ri, *rj, *du_prev, *dv_prev;
    const float *du, *dv, *dw;        // last primal iterate
    const float *I1, *I2, *Zold;      // finest pyramid planes
    const float *du_l, *dv_l;         // final flow (pixels)
    unsigned rows, cols;
    float lambda_i, lambda_d;
};

// end-point error between two flows of one pair
struct EpeArgs {
    const float *ax, *ay, *az, *bx, *by, *bz, *zold;
    unsigned pixels;
};

enum class LearnKernel { GatherInputs, Inject, Labels, Metrics, Epe };

constexpr unsigned kLearnBlock = 256;
inline unsigned learn_blocks(unsigned pixels) { return (pixels + kLearnBlock - 1) / kLearnBlock; }

// partial sums per block of the Metrics reduction: energy, warped residual, pixels with
// frame-1 depth, zero-motion residual
constexpr int kMetricPartials = 4;

// batched launch over a BatchTable of LearnArgs / MetricArgs / EpeArgs; the reductions write
// per-block partial sums (Metrics: kMetricPartials floats, Epe: 2) at partial + stride * block
hipError_t launch_learn(LearnKernel k, const BatchTable &t, unsigned total_blocks, float *partial,
                        hipStream_t s);

}  // namespace pdf

namespace pdf {
namespace cnn {
// the network's forward pass on a caller's stream (weights re-cast from the FP32 master copy
// first, so captured replays follow later weight updates); output at net_output()
int net_forward_on(pdf_net *n, const float *x_dev, unsigned nb, hipStream_t s);
const float *net_output(const pdf_net *n);
void net_dims(const pdf_net *n, unsigned *h, unsigned *w, unsigned *max_batch);
}  // namespace cnn
}  // namespace pdf
