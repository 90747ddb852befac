// Learned warm start glue and per-pair quality reductions for the batched solver.  No
// counterpart in the reference.
//
//  - GatherInputs: a pair's network-size pyramid level (I1, Z1, I2, Z2; column-major) into
//    the network's NHWC input, normalised I/255 - 0.5 and Z/4;
//  - Inject: the network's prediction (du, dv in pixels of that level, 10*dw) into the pair's
//    warm-start planes, which the WarmInit set-up kernel turns into du_prev, dv_prev, dw;
//  - Labels: the finest flow of a solved pair averaged over each network pixel's block
//    (du, dv rescaled to that level's pixels, dw times 10), the training target;
//  - Metrics: the finest level's energy on the last primal iterate,
//      E = sum |dct + dcu du + dcv dv| + mu_uv |ddt + ddu du + ddv dv - dw|
//          + lambda_i (|grad_r(du + du_prev)| + |grad_r(dv + dv_prev)|) + lambda_d |grad_r dw|,
//    the L1 data terms of the prox (pdflow_cudalib.cu:906-923) and the depth dual
//    (.cu:836-846), the TV terms of the duals (.cu:849-885) with grad_r the r-weighted forward
//    differences of .cu:982-1013, and the warped-intensity residual |I2(x + u) - I1(x)| over
//    pixels with frame-1 depth (bilinear warp of .cu:654-682), and the zero-motion residual
//    |I2(x) - I1(x)| over the same pixels (the intensity change the flow has to explain);
//  - Epe: the 3-D end-point error between two flows over pixels with frame-1 depth.
// Reductions: wave64 butterfly (__shfl_xor) then the block's four waves through LDS; one
// partial per block, summed per pair on the host in double precision.
#include <hip/hip_runtime.h>

#include "pdflow_learn.hpp"
#include "pdflow_pixel.hpp"

namespace pdf {
namespace {

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sums of N values over the block (256 threads); result valid in thread 0
template <int N>
__device__ __forceinline__ void block_sum(float (&v)[N], float *out)
{
    __shared__ float s[4][N];
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < N; i++) v[i] = wave_sum(v[i]);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < N; i++) s[wave][i] = v[i];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < N; i++) out[i] = (s[0][i] + s[1][i]) + (s[2][i] + s[3][i]);
}

__global__ void __launch_bounds__(kLearnBlock) kb_gather_inputs(BatchTable t)
{
    const BatchItem it = batch_item(t);
    const LearnArgs &a = static_cast<const LearnArgs *>(t.args)[it.item];
    const unsigned i = it.block * kLearnBlock + threadIdx.x;     // row-major pixel
    if (i >= a.rows * a.cols) return;
    const unsigned v = i / a.cols, u = i % a.cols, c = v + u * a.rows;
    const float4 x = make_float4(gptr(a.I1)[c] * (1.f / 255.f) - 0.5f, gptr(a.Z1)[c] * 0.25f,
                                 gptr(a.I2)[c] * (1.f / 255.f) - 0.5f, gptr(a.Z2)[c] * 0.25f);
    reinterpret_cast<float4 *>(a.x)[i] = x;
}

__global__ void __launch_bounds__(kLearnBlock) kb_inject(BatchTable t)
{
    const BatchItem it = batch_item(t);
    const LearnArgs &a = static_cast<const LearnArgs *>(t.args)[it.item];
    const unsigned P = a.rows * a.cols;
    const unsigned i = it.block * kLearnBlock + threadIdx.x;
    if (i >= P) return;
    const unsigned v = i / a.cols, u = i % a.cols, c = v + u * a.rows;
    gptr(a.warm)[c] = gptr(a.y)[3 * i];
    gptr(a.warm)[P + c] = gptr(a.y)[3 * i + 1];
    gptr(a.warm)[2 * P + c] = gptr(a.y)[3 * i + 2] / 10.f;
}

__global__ void __launch_bounds__(kLearnBlock) kb_labels(BatchTable t)
{
    const BatchItem it = batch_item(t);
    const LearnArgs &a = static_cast<const LearnArgs *>(t.args)[it.item];
    const unsigned i = it.block * kLearnBlock + threadIdx.x;
    if (i >= a.rows * a.cols) return;
    const unsigned v = i / a.cols, u = i % a.cols;
    const unsigned s = a.frows / a.rows;
    float su = 0.f, sv = 0.f, sw = 0.f;
    for (unsigned du = 0; du < s; du++)
        for (unsigned dv = 0; dv < s; dv++) {
            const unsigned c = (v * s + dv) + (u * s + du) * a.frows;
            su += gptr(a.du)[c];
            sv += gptr(a.dv)[c];
            sw += gptr(a.dw)[c];
        }
    const float inv = 1.f / (float)(s * s);
    gptr(a.t)[3 * i] = su * inv / (float)s;
    gptr(a.t)[3 * i + 1] = sv * inv / (float)s;
    gptr(a.t)[3 * i + 2] = sw * inv * 10.f;
    gptr(a.m)[i] = 1.f;
}

__global__ void __launch_bounds__(kLearnBlock) kb_metrics(BatchTable t, float *partial)
{
    // blocks in XCD-contiguous order (neighbouring columns meet in one L2); each block's partial
    // sums go to its place in that order, so the host's per-pair sums are unchanged
    const unsigned lb = xcd_swizzle(blockIdx.x, gridDim.x);
    const BatchItem it = batch_item(t, lb);
    const MetricArgs &a = static_cast<const MetricArgs *>(t.args)[it.item];
    const unsigned rows = a.rows, cols = a.cols;
    const unsigned idx = it.block * kLearnBlock + threadIdx.x;
    float acc[kMetricPartials] = {0.f, 0.f, 0.f, 0.f};
    if (idx < rows * cols) {
        const unsigned v = idx % rows, u = idx / rows;
        const float du = gptr(a.du)[idx], dv = gptr(a.dv)[idx], dw = gptr(a.dw)[idx];
        const float ub = du + gptr(a.du_prev)[idx], vb = dv + gptr(a.dv_prev)[idx];
        const float ri = gptr(a.ri)[idx], rj = gptr(a.rj)[idx];
        float gu1 = 0.f, gv1 = 0.f, gw1 = 0.f, gu2 = 0.f, gv2 = 0.f, gw2 = 0.f;
        if (u + 1 < cols) {
            const unsigned n = idx + rows;
            gu1 = ri * ((gptr(a.du)[n] + gptr(a.du_prev)[n]) - ub);
            gv1 = ri * ((gptr(a.dv)[n] + gptr(a.dv_prev)[n]) - vb);
            gw1 = ri * (gptr(a.dw)[n] - dw);
        }
        if (v + 1 < rows) {
            const unsigned n = idx + 1;
            gu2 = rj * ((gptr(a.du)[n] + gptr(a.du_prev)[n]) - ub);
            gv2 = rj * ((gptr(a.dv)[n] + gptr(a.dv_prev)[n]) - vb);
            gw2 = rj * (gptr(a.dw)[n] - dw);
        }
        const float data_i = fabsf(gptr(a.dct)[idx] + gptr(a.dcu)[idx] * du + gptr(a.dcv)[idx] * dv);
        const float data_d = gptr(a.mu_uv)[idx] * fabsf(gptr(a.ddt)[idx] + gptr(a.ddu)[idx] * du + gptr(a.ddv)[idx] * dv - dw);
        const float tv = a.lambda_i * (sqrtf(gu1 * gu1 + gu2 * gu2) + sqrtf(gv1 * gv1 + gv2 * gv2)) +
                         a.lambda_d * sqrtf(gw1 * gw1 + gw2 * gw2);
        acc[0] = data_i + data_d + tv;
        if (gptr(a.Zold)[idx] > 0.f) {
            const Tap tp = make_tap(rows, cols, (float)u + gptr(a.du_l)[idx], (float)v + gptr(a.dv_l)[idx]);
            const float i1 = gptr(a.I1)[idx];
            acc[1] = fabsf(bilinear(a.I2, rows, tp) - i1);
            acc[2] = 1.f;
            acc[3] = fabsf(gptr(a.I2)[idx] - i1);
        }
    }
    block_sum<kMetricPartials>(acc, partial + kMetricPartials * (size_t)lb);
}

__global__ void __launch_bounds__(kLearnBlock) kb_epe(BatchTable t, float *partial)
{
    const BatchItem it = batch_item(t);
    const EpeArgs &a = static_cast<const EpeArgs *>(t.args)[it.item];
    const unsigned i = it.block * kLearnBlock + threadIdx.x;
    float acc[2] = {0.f, 0.f};
    if (i < a.pixels && gptr(a.zold)[i] > 0.f) {
        const float ex = gptr(a.ax)[i] - gptr(a.bx)[i], ey = gptr(a.ay)[i] - gptr(a.by)[i], ez = gptr(a.az)[i] - gptr(a.bz)[i];
        acc[0] = sqrtf(ex * ex + ey * ey + ez * ez);
        acc[1] = 1.f;
    }
    block_sum<2>(acc, partial + 2 * (size_t)blockIdx.x);
}

}  // namespace

hipError_t launch_learn(LearnKernel k, const BatchTable &t, unsigned total_blocks, float *partial, hipStream_t s)
{
    const dim3 grid(total_blocks), block(kLearnBlock);
    switch (k) {
    case LearnKernel::GatherInputs: hipLaunchKernelGGL(kb_gather_inputs, grid, block, 0, s, t); break;
    case LearnKernel::Inject: hipLaunchKernelGGL(kb_inject, grid, block, 0, s, t); break;
    case LearnKernel::Labels: hipLaunchKernelGGL(kb_labels, grid, block, 0, s, t); break;
    case LearnKernel::Metrics: hipLaunchKernelGGL(kb_metrics, grid, block, 0, s, t, partial); break;
    case LearnKernel::Epe: hipLaunchKernelGGL(kb_epe, grid, block, 0, s, t, partial); break;
    }
    return hipGetLastError();
}

}  // namespace pdf
