// gfx950 kernels of the PD-Flow plain solve.
//
// Per-pixel arithmetic follows the reference's device bodies (pdflow_cudalib.cu:331-1167)
// operation for operation: FP32, evaluated in the order written there, compiled with
// -ffp-contract=off, IEEE division and square root.  rhypotf(a,b) is 1/sqrtf(a*a+b*b) (the
// reference's own comment at .cu:761), powf(x,2) is x*x, fdividef is IEEE division.
//
// Launch shapes are this library's own: 256-thread workgroups (four wave64s) over a 1-D
// pixel range for the setup stages, and 32x8-pixel LDS tiles for the fused primal-dual
// iteration.  Nothing here uses MFMA: every stage is a memory-bound stencil.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pdflow_internal.hpp"
#include "pdflow_pixel.hpp"

namespace pdf {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ unsigned pixel_index() { return blockIdx.x * kBlock + threadIdx.x; }

inline dim3 grid_for(unsigned n) { return dim3((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------------ K1 pyramid level
// computePyramidLevel (.cu:331-434).  One thread per output pixel; the 5x5 mask comes in
// the kernel arguments, so no block barrier is needed.
__global__ void __launch_bounds__(kBlock) k_pyramid(PyrArgs a)
{
    const unsigned index = pixel_index();
    const unsigned rows_i = a.rows_i, cols_i = a.cols_i;
    if (index >= rows_i * cols_i) return;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float max_depth_dif = 0.1f;
    float depth;
    if (a.level == 0) {
        a.col[index] = a.colour_wf[index];
        depth = a.depth_wf[index];
        a.dep[index] = depth;
    } else {
        float sumd = 0.f, sumc = 0.f, wd = 0.f, wc = 0.f;
        const float dcen = a.depp[2 * v + 4 * u * rows_i];
        const bool inner = (v > 0) && (v < rows_i - 1) && (u > 0) && (u < cols_i - 1);
#pragma unroll
        for (int k = -2; k < 3; k++)
#pragma unroll
            for (int l = -2; l < 3; l++) {
                const int vv = 2 * (int)v + k, uu = 2 * (int)u + l;
                if (!inner && !(vv >= 0 && vv < 2 * (int)rows_i && uu >= 0 && uu < 2 * (int)cols_i))
                    continue;
                const unsigned ip = (unsigned)vv + (unsigned)uu * 2 * rows_i;
                const float g = a.g[12 + k + 5 * l];
                sumc += g * a.colp[ip];
                if (!inner) wc += g;
                const float d = a.depp[ip];
                if ((d > 0.f) && (fabsf(d - dcen) < max_depth_dif)) {
                    const float w = g * (max_depth_dif - fabsf(d - dcen));
                    wd += w;
                    sumd += w * d;
                }
            }
        depth = (sumd > 0.f) ? sumd / wd : 0.f;
        a.dep[index] = depth;
        a.col[index] = inner ? sumc : sumc / wc;
    }
    const float disp_u = 0.5f * (float)(cols_i - 1);
    const float disp_v = 0.5f * (float)(rows_i - 1);
    a.xx[index] = ((float)u - disp_u) * depth * a.inv_f_i;
    a.yy[index] = ((float)v - disp_v) * depth * a.inv_f_i;
}

// ------------------------------------------------------------------ K2 assign zeros
__global__ void __launch_bounds__(kBlock) k_assign_zeros(LevelArgs a)
{
    // .cu:439-455
    const unsigned index = pixel_index();
    if (index >= a.rows * a.cols) return;
    constexpr int ids[] = {PDF_DU_UPSAMP, PDF_DV_UPSAMP, PDF_DW_UPSAMP, PDF_PD_UPSAMP,
                           PDF_PUU_UPSAMP, PDF_PUV_UPSAMP, PDF_PVU_UPSAMP, PDF_PVV_UPSAMP,
                           PDF_PWU_UPSAMP, PDF_PWV_UPSAMP, PDF_DU_PREV, PDF_DV_PREV,
                           PDF_DU_NEW, PDF_DV_NEW, PDF_DW_NEW, PDF_PD, PDF_PUU, PDF_PUV,
                           PDF_PVU, PDF_PVV, PDF_PWU, PDF_PWV, PDF_DU_ACC, PDF_DV_ACC,
                           PDF_DW_ACC};
#pragma unroll
    for (int k = 0; k < (int)(sizeof(ids) / sizeof(ids[0])); k++) a.f[ids[k]][index] = 0.f;
}

// ------------------------------------------------------------------ K3/K4 upsampling
__constant__ int kUps[10] = {PDF_DU_UPSAMP, PDF_DV_UPSAMP, PDF_DW_UPSAMP, PDF_PD_UPSAMP,
                             PDF_PUU_UPSAMP, PDF_PUV_UPSAMP, PDF_PVU_UPSAMP, PDF_PVV_UPSAMP,
                             PDF_PWU_UPSAMP, PDF_PWV_UPSAMP};

__global__ void __launch_bounds__(kBlock) k_upsample_copy(LevelArgs a)
{
    // .cu:460-477: coarse (v,u) -> fine (2v,2u); du, dv doubled (pixel units)
    const unsigned index = pixel_index();
    const unsigned rows_i = a.rows;
    if (index >= rows_i * a.cols / 4) return;
    const unsigned v = index % (rows_i / 2);
    const unsigned u = 2 * index / rows_i;
    const unsigned big = 2 * (v + u * rows_i);
    a.f[PDF_DU_UPSAMP][big] = 2.f * a.f[PDF_DU_L][index];
    a.f[PDF_DV_UPSAMP][big] = 2.f * a.f[PDF_DV_L][index];
    a.f[PDF_DW_UPSAMP][big] = a.f[PDF_DW_L][index];
    a.f[PDF_PD_UPSAMP][big] = a.f[PDF_PD_L][index];
    a.f[PDF_PUU_UPSAMP][big] = a.f[PDF_PUU_L][index];
    a.f[PDF_PUV_UPSAMP][big] = a.f[PDF_PUV_L][index];
    a.f[PDF_PVU_UPSAMP][big] = a.f[PDF_PVU_L][index];
    a.f[PDF_PVV_UPSAMP][big] = a.f[PDF_PVV_L][index];
    a.f[PDF_PWU_UPSAMP][big] = a.f[PDF_PWU_L][index];
    a.f[PDF_PWV_UPSAMP][big] = a.f[PDF_PWV_L][index];
}

__global__ void __launch_bounds__(kBlock) k_upsample_filter(LevelArgs a)
{
    // .cu:479-569: 5x5 filter with 4g over the zero-inserted planes; at the border the
    // weight is 1 - sum of the out-of-range g
    const unsigned index = pixel_index();
    const unsigned rows_i = a.rows, cols_i = a.cols;
    if (index >= rows_i * cols_i) return;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float *src[10];
#pragma unroll
    for (int q = 0; q < 10; q++) src[q] = a.f[kUps[q]];
    float acc[10];
#pragma unroll
    for (int q = 0; q < 10; q++) acc[q] = 0.f;
    if ((v > 1) && (v < rows_i - 2) && (u > 1) && (u < cols_i - 2)) {
#pragma unroll
        for (int k = -2; k < 3; k++)
#pragma unroll
            for (int l = -2; l < 3; l++) {
                const unsigned ii = v + k + (u + l) * rows_i;
                const float g = 4.f * a.g[12 + k + 5 * l];
#pragma unroll
                for (int q = 0; q < 10; q++) acc[q] += g * src[q][ii];
            }
    } else {
        float aw = 1.f;
#pragma unroll
        for (int k = -2; k < 3; k++)
#pragma unroll
            for (int l = -2; l < 3; l++) {
                const int vv = (int)v + k, uu = (int)u + l;
                const float g = 4.f * a.g[12 + k + 5 * l];
                if (vv < 0 || vv >= (int)rows_i || uu < 0 || uu >= (int)cols_i) {
                    aw -= 0.25f * g;
                    continue;
                }
                const unsigned ii = (unsigned)vv + (unsigned)uu * rows_i;
#pragma unroll
                for (int q = 0; q < 10; q++) acc[q] += g * src[q][ii];
            }
        const float inv = 1.f / aw;
#pragma unroll
        for (int q = 0; q < 10; q++) acc[q] *= inv;
    }
    a.f[PDF_DU_PREV][index] = acc[0];
    a.f[PDF_DV_PREV][index] = acc[1];
    a.f[PDF_DW_NEW][index] = acc[2];
    a.f[PDF_PD][index] = acc[3];
    a.f[PDF_PUU][index] = acc[4];
    a.f[PDF_PUV][index] = acc[5];
    a.f[PDF_PVU][index] = acc[6];
    a.f[PDF_PVV][index] = acc[7];
    a.f[PDF_PWU][index] = acc[8];
    a.f[PDF_PWV][index] = acc[9];
    a.f[PDF_DW_ACC][index] = acc[2];
}

// ------------------------------------------------------------------ K5 Rij
__global__ void __launch_bounds__(kBlock) k_rij(LevelArgs a)
{
    // .cu:725-779
    const unsigned index = pixel_index();
    const unsigned rows_i = a.rows, cols_i = a.cols;
    if (index >= rows_i * cols_i) return;
    const unsigned v = index % rows_i, u = index / rows_i;
    float dxu = 0.f, dzu = 0.f, dxu2 = 0.f, dzu2 = 0.f;
    float dyv = 0.f, dzv = 0.f, dyv2 = 0.f, dzv2 = 0.f;
    if (u != cols_i - 1) {
        dxu = a.Xold[index + rows_i] - a.Xold[index];
        dzu = a.Zold[index + rows_i] - a.Zold[index];
        dxu2 = a.X[index + rows_i] - a.X[index];
        dzu2 = a.Z[index + rows_i] - a.Z[index];
    }
    if (v != rows_i - 1) {
        dyv = a.Yold[index + 1] - a.Yold[index];
        dzv = a.Zold[index + 1] - a.Zold[index];
        dyv2 = a.Y[index + 1] - a.Y[index];
        dzv2 = a.Z[index + 1] - a.Z[index];
    }
    a.f[PDF_RI][index] = (fabsf(dxu) + fabsf(dzu) > 0.f) ? 2.f * rhyp(dxu, dzu) : 1.f;
    a.f[PDF_RJ][index] = (fabsf(dyv) + fabsf(dzv) > 0.f) ? 2.f * rhyp(dyv, dzv) : 1.f;
    a.f[PDF_RI_2][index] = (fabsf(dxu2) + fabsf(dzu2) > 0.f) ? 2.f * rhyp(dxu2, dzu2) : 1.f;
    a.f[PDF_RJ_2][index] = (fabsf(dyv2) + fabsf(dzv2) > 0.f) ? 2.f * rhyp(dyv2, dzv2) : 1.f;
}

// ------------------------------------------------------------------ K6 image gradients
__global__ void __launch_bounds__(kBlock) k_image_gradients(LevelArgs a)
{
    // .cu:574-627
    const unsigned index = pixel_index();
    const unsigned rows_i = a.rows, cols_i = a.cols, r = rows_i;
    if (index >= rows_i * cols_i) return;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float *I = a.I, *Z = a.Z, *ri2 = a.f[PDF_RI_2], *rj2 = a.f[PDF_RJ_2];
    float dcu, ddu, dcv, ddv;
    if (u == 0) {
        dcu = I[index + r] - I[index];
        ddu = Z[index + r] - Z[index];
    } else if (u == cols_i - 1) {
        dcu = I[index] - I[index - r];
        ddu = Z[index] - Z[index - r];
    } else {
        const float rn = ri2[index], rp = ri2[index - r];
        dcu = wdiff(rn, rp, I[index + r], I[index], I[index - r]);
        ddu = (Z[index] > 0.f) ? wdiff(rn, rp, Z[index + r], Z[index], Z[index - r]) : 0.f;
    }
    if (v == 0) {
        dcv = I[index + 1] - I[index];
        ddv = Z[index + 1] - Z[index];
    } else if (v == rows_i - 1) {
        dcv = I[index] - I[index - 1];
        ddv = Z[index] - Z[index - 1];
    } else {
        const float rn = rj2[index], rp = rj2[index - 1];
        dcv = wdiff(rn, rp, I[index + 1], I[index], I[index - 1]);
        ddv = (Z[index] > 0.f) ? wdiff(rn, rp, Z[index + 1], Z[index], Z[index - 1]) : 0.f;
    }
    a.f[PDF_DCU_AUX][index] = dcu;
    a.f[PDF_DDU_AUX][index] = ddu;
    a.f[PDF_DCV_AUX][index] = dcv;
    a.f[PDF_DDV_AUX][index] = ddv;
}

// ------------------------------------------------------------------ K7 warping
__global__ void __launch_bounds__(kBlock) k_warping(LevelArgs a)
{
    // .cu:629-652
    const unsigned index = pixel_index();
    const unsigned rows_i = a.rows, cols_i = a.cols;
    if (index >= rows_i * cols_i) return;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float iu = (float)u + a.f[PDF_DU_PREV][index];
    const float iv = (float)v + a.f[PDF_DV_PREV][index];
    const Tap t = make_tap(rows_i, cols_i, iu, iv);
    a.f[PDF_DCT][index] = bilinear(a.I, rows_i, t) - a.Iold[index];
    a.f[PDF_DCU][index] = bilinear(a.f[PDF_DCU_AUX], rows_i, t);
    a.f[PDF_DCV][index] = bilinear(a.f[PDF_DCV_AUX], rows_i, t);
    const float wd = bilinear_depth(a.Z, rows_i, t);
    a.f[PDF_DDT][index] = (wd > 0.f) ? wd - a.Zold[index] : 0.f;
    a.f[PDF_DDU][index] = bilinear(a.f[PDF_DDU_AUX], rows_i, t);
    a.f[PDF_DDV][index] = bilinear(a.f[PDF_DDV_AUX], rows_i, t);
}

// ------------------------------------------------------------------ K8 mu and steps
__global__ void __launch_bounds__(kBlock) k_mu_steps(LevelArgs a)
{
    // computeMu + computeStepSizes (.cu:781-808)
    const unsigned index = pixel_index();
    const unsigned rows_i = a.rows;
    if (index >= rows_i * a.cols) return;
    const float ddu = a.f[PDF_DDU][index], ddv = a.f[PDF_DDV][index], ddt = a.f[PDF_DDT][index];
    const float mu_uv = a.mu / (1.f + 1000.f * (ddu * ddu + ddv * ddv + ddt * ddt));
    a.f[PDF_MU_UV][index] = mu_uv;
    const float ri = a.f[PDF_RI][index], rj = a.f[PDF_RJ][index];
    a.f[PDF_SIGMA_PD][index] = sigma_pd_of(mu_uv, ddu, ddv);
    a.f[PDF_SIGMA_PUVX][index] = sigma_p_of(a.lambda_i, ri);
    a.f[PDF_SIGMA_PUVY][index] = sigma_p_of(a.lambda_i, rj);
    a.f[PDF_SIGMA_PWX][index] = 0.5f / (ri * a.lambda_d + 1e-10f);
    a.f[PDF_SIGMA_PWY][index] = 0.5f / (rj * a.lambda_d + 1e-10f);
    const unsigned v = index % rows_i, u = index / rows_i;
    float acu_r = ri + rj;
    if (u > 0) acu_r += a.f[PDF_RI][index - rows_i];
    if (v > 0) acu_r += a.f[PDF_RJ][index - 1];
    a.f[PDF_TAU_U][index] = tau_uv_of(mu_uv, ddu, a.lambda_i, acu_r);
    a.f[PDF_TAU_V][index] = tau_uv_of(mu_uv, ddv, a.lambda_i, acu_r);
    a.f[PDF_TAU_W][index] = tau_w_of(mu_uv, a.lambda_d, acu_r);
}

// ------------------------------------------------------------------ K9-K12 (unfused)
__global__ void __launch_bounds__(kBlock) k_gradient(LevelArgs a)
{
    // .cu:982-1013
    const unsigned index = pixel_index();
    const unsigned rows_i = a.rows, cols_i = a.cols;
    if (index >= rows_i * cols_i) return;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float *ua = a.f[PDF_DU_ACC], *va = a.f[PDF_DV_ACC], *wa = a.f[PDF_DW_ACC];
    const float *up = a.f[PDF_DU_PREV], *vp = a.f[PDF_DV_PREV];
    const float ri = a.f[PDF_RI][index], rj = a.f[PDF_RJ][index];
    const float u0 = ua[index] + up[index], v0 = va[index] + vp[index], w0 = wa[index];
    if (u == cols_i - 1) {
        a.f[PDF_GRADU1][index] = 0.f; a.f[PDF_GRADV1][index] = 0.f; a.f[PDF_GRADW1][index] = 0.f;
    } else {
        const unsigned n = index + rows_i;
        a.f[PDF_GRADU1][index] = ri * ((ua[n] + up[n]) - u0);
        a.f[PDF_GRADV1][index] = ri * ((va[n] + vp[n]) - v0);
        a.f[PDF_GRADW1][index] = ri * (wa[n] - w0);
    }
    if (v == rows_i - 1) {
        a.f[PDF_GRADU2][index] = 0.f; a.f[PDF_GRADV2][index] = 0.f; a.f[PDF_GRADW2][index] = 0.f;
    } else {
        const unsigned n = index + 1;
        a.f[PDF_GRADU2][index] = rj * ((ua[n] + up[n]) - u0);
        a.f[PDF_GRADV2][index] = rj * ((va[n] + vp[n]) - v0);
        a.f[PDF_GRADW2][index] = rj * (wa[n] - w0);
    }
}

__global__ void __launch_bounds__(kBlock) k_dual(LevelArgs a)
{
    // .cu:813-887
    const unsigned index = pixel_index();
    if (index >= a.rows * a.cols) return;
    const float li = a.lambda_i, ld = a.lambda_d;
    float pd = a.f[PDF_PD][index], puu = a.f[PDF_PUU][index], puv = a.f[PDF_PUV][index];
    float pvu = a.f[PDF_PVU][index], pvv = a.f[PDF_PVV][index];
    float pwu = a.f[PDF_PWU][index], pwv = a.f[PDF_PWV][index];
    pd += a.f[PDF_SIGMA_PD][index] * a.f[PDF_MU_UV][index] *
          (-a.f[PDF_DW_ACC][index] + a.f[PDF_DDT][index] + a.f[PDF_DDU][index] * a.f[PDF_DU_ACC][index] +
           a.f[PDF_DDV][index] * a.f[PDF_DV_ACC][index]);
    puu += a.f[PDF_SIGMA_PUVX][index] * li * a.f[PDF_GRADU1][index];
    puv += a.f[PDF_SIGMA_PUVY][index] * li * a.f[PDF_GRADU2][index];
    pvu += a.f[PDF_SIGMA_PUVX][index] * li * a.f[PDF_GRADV1][index];
    pvv += a.f[PDF_SIGMA_PUVY][index] * li * a.f[PDF_GRADV2][index];
    pwu += a.f[PDF_SIGMA_PWX][index] * ld * a.f[PDF_GRADW1][index];
    pwv += a.f[PDF_SIGMA_PWY][index] * ld * a.f[PDF_GRADW2][index];
    a.f[PDF_PD][index] = clamp_pd(pd);
    project(puu, puv);
    project(pvu, pvv);
    project(pwu, pwv);
    a.f[PDF_PUU][index] = puu; a.f[PDF_PUV][index] = puv;
    a.f[PDF_PVU][index] = pvu; a.f[PDF_PVV][index] = pvv;
    a.f[PDF_PWU][index] = pwu; a.f[PDF_PWV][index] = pwv;
}

__global__ void __launch_bounds__(kBlock) k_divergence(LevelArgs a)
{
    // .cu:935-980
    const unsigned index = pixel_index();
    const unsigned rows_i = a.rows, cols_i = a.cols;
    if (index >= rows_i * cols_i) return;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float *ri = a.f[PDF_RI], *rj = a.f[PDF_RJ];
    const bool uf = (u == 0), ul = (u == cols_i - 1), vf = (v == 0), vl = (v == rows_i - 1);
    const unsigned iu = uf ? index : index - rows_i;   // never read when uf
    const unsigned iv = vf ? index : index - 1;
    float du = div_axis(ri[index], a.f[PDF_PUU][index], ri[iu], a.f[PDF_PUU][iu], uf, ul);
    float dv = div_axis(ri[index], a.f[PDF_PVU][index], ri[iu], a.f[PDF_PVU][iu], uf, ul);
    float dw = div_axis(ri[index], a.f[PDF_PWU][index], ri[iu], a.f[PDF_PWU][iu], uf, ul);
    du += div_axis(rj[index], a.f[PDF_PUV][index], rj[iv], a.f[PDF_PUV][iv], vf, vl);
    dv += div_axis(rj[index], a.f[PDF_PVV][index], rj[iv], a.f[PDF_PVV][iv], vf, vl);
    dw += div_axis(rj[index], a.f[PDF_PWV][index], rj[iv], a.f[PDF_PWV][iv], vf, vl);
    a.f[PDF_DIVPU][index] = du;
    a.f[PDF_DIVPV][index] = dv;
    a.f[PDF_DIVPW][index] = dw;
}

__global__ void __launch_bounds__(kBlock) k_primal(LevelArgs a)
{
    // .cu:889-933
    const unsigned index = pixel_index();
    if (index >= a.rows * a.cols) return;
    const float du0 = a.f[PDF_DU_NEW][index], dv0 = a.f[PDF_DV_NEW][index], dw0 = a.f[PDF_DW_NEW][index];
    const Primal p = primal_update(du0, dv0, dw0, a.f[PDF_TAU_U][index], a.f[PDF_TAU_V][index],
                                   a.f[PDF_TAU_W][index], a.f[PDF_MU_UV][index], a.f[PDF_PD][index],
                                   a.f[PDF_DDU][index], a.f[PDF_DDV][index], a.f[PDF_DIVPU][index],
                                   a.f[PDF_DIVPV][index], a.f[PDF_DIVPW][index], a.f[PDF_DCT][index],
                                   a.f[PDF_DCU][index], a.f[PDF_DCV][index], a.lambda_i, a.lambda_d);
    a.f[PDF_DU_ACC][index] = 2.f * p.du - du0;
    a.f[PDF_DV_ACC][index] = 2.f * p.dv - dv0;
    a.f[PDF_DW_ACC][index] = 2.f * p.dw - dw0;
    a.f[PDF_DU_NEW][index] = p.du;
    a.f[PDF_DV_NEW][index] = p.dv;
    a.f[PDF_DW_NEW][index] = p.dw;
}

// ------------------------------------------------------------------ K13/K14 filter
__global__ void __launch_bounds__(kBlock) k_saturate(LevelArgs a)
{
    // .cu:1018-1038
    const unsigned index = pixel_index();
    if (index >= a.rows * a.cols) return;
    float du = a.f[PDF_DU_NEW][index], dv = a.f[PDF_DV_NEW][index], dw = a.f[PDF_DW_NEW][index];
    if (du > 1.f) du = 1.f; else if (du < -1.f) du = -1.f;
    if (dv > 1.f) dv = 1.f; else if (dv < -1.f) dv = -1.f;
    if (a.Zold[index] == 0.f) dw = 0.f;
    a.f[PDF_DU_NEW][index] = du + a.f[PDF_DU_PREV][index];
    a.f[PDF_DV_NEW][index] = dv + a.f[PDF_DV_PREV][index];
    a.f[PDF_DW_NEW][index] = dw;
}

__global__ void __launch_bounds__(kBlock) k_filter(LevelArgs a)
{
    // .cu:1040-1148
    const unsigned index = pixel_index();
    const unsigned rows_i = a.rows, cols_i = a.cols;
    if (index >= rows_i * cols_i) return;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float *zo = a.Zold;
    const float d0 = zo[index];
    if (d0 > 0.f) {
        float ku[9], kv[9], kw[9], wu[9], wv[9], ww[9];
        int n = 0;
#pragma unroll
        for (int s = 0; s < 9; s++) {
            ku[s] = kv[s] = kw[s] = __builtin_inff();
            wu[s] = wv[s] = ww[s] = 0.f;
        }
        // taps in the reference's order: k (v offset) outer, l (u offset) inner
#pragma unroll
        for (int k = -1; k < 2; k++)
#pragma unroll
            for (int l = -1; l < 2; l++) {
                const int r = (int)v + k, c = (int)u + l;
                const bool ok = !(r < 0 || r >= (int)rows_i || c < 0 || c >= (int)cols_i);
                const unsigned il = ok ? index + l * (int)rows_i + k : index;
                const float dz = d0 - zo[il], dt = a.f[PDF_DDT][il];
                const float pres = 1.f / (1.f + 5.f * (dz * dz) + 10.f * (dt * dt));
                const float fu = a.f[PDF_DU_NEW][il], fv = a.f[PDF_DV_NEW][il], fw = a.f[PDF_DW_NEW][il];
                // compacted write position n (taps out of range are skipped)
#pragma unroll
                for (int s = 0; s < 9; s++)
                    if (ok && s == n) {
                        ku[s] = fu; kv[s] = fv; kw[s] = fw;
                        wu[s] = pres; wv[s] = pres; ww[s] = pres;
                    }
                n += ok ? 1 : 0;
            }
        sort9(ku, wu);
        sort9(kv, wv);
        sort9(kw, ww);
        float cu[9], cv[9], cw[9];
        cu[0] = wu[0]; cv[0] = wv[0]; cw[0] = ww[0];
#pragma unroll
        for (int i = 1; i < 9; i++) {
            cu[i] = cu[i - 1] + wu[i];
            cv[i] = cv[i - 1] + wv[i];
            cw[i] = cw[i - 1] + ww[i];
        }
        // total of u's weights = cu[n-1]; adding the zero weights of unused slots keeps it
        const float med = 0.5f * cu[8];
        a.f[PDF_DU_L][index] = wmedian_pick(ku, cu, med);
        a.f[PDF_DV_L][index] = wmedian_pick(kv, cv, med);
        a.f[PDF_DW_L][index] = wmedian_pick(kw, cw, med);
    } else {
        a.f[PDF_DU_L][index] = a.f[PDF_DU_NEW][index];
        a.f[PDF_DV_L][index] = a.f[PDF_DV_NEW][index];
        a.f[PDF_DW_L][index] = a.f[PDF_DW_NEW][index];
    }
    a.f[PDF_PD_L][index] = a.f[PDF_PD][index];
    a.f[PDF_PUU_L][index] = a.f[PDF_PUU][index];
    a.f[PDF_PUV_L][index] = a.f[PDF_PUV][index];
    a.f[PDF_PVU_L][index] = a.f[PDF_PVU][index];
    a.f[PDF_PVV_L][index] = a.f[PDF_PVV][index];
    a.f[PDF_PWU_L][index] = a.f[PDF_PWU][index];
    a.f[PDF_PWV_L][index] = a.f[PDF_PWV][index];
}

// ------------------------------------------------------------------ K15 motion field
__global__ void __launch_bounds__(kBlock) k_motion_field(LevelArgs a)
{
    // .cu:1150-1167
    const unsigned index = pixel_index();
    if (index >= a.rows * a.cols) return;
    const float z = a.Zold[index];
    if (z > 0.f) {
        const float dul = a.f[PDF_DU_L][index], dvl = a.f[PDF_DV_L][index], dwl = a.f[PDF_DW_L][index];
        a.f[PDF_DX][index] = dwl;
        a.f[PDF_DY][index] = z * dul * a.inv_f + dwl * a.Xold[index] / z;
        a.f[PDF_DZ][index] = z * dvl * a.inv_f + dwl * a.Yold[index] / z;
    } else {
        a.f[PDF_DX][index] = 0.f;
        a.f[PDF_DY][index] = 0.f;
        a.f[PDF_DZ][index] = 0.f;
    }
}

}  // namespace

#define PDF_LAUNCH(kernel, n, args, stream)                                  \
    do {                                                                     \
        if ((n) == 0) return hipSuccess;                                     \
        hipLaunchKernelGGL(kernel, grid_for(n), dim3(kBlock), 0, stream, args); \
        return hipGetLastError();                                            \
    } while (0)

hipError_t launch_pyramid(const PyrArgs &a, hipStream_t s) { PDF_LAUNCH(k_pyramid, a.rows_i * a.cols_i, a, s); }
hipError_t launch_assign_zeros(const LevelArgs &a, hipStream_t s) { PDF_LAUNCH(k_assign_zeros, a.rows * a.cols, a, s); }
hipError_t launch_upsample(const LevelArgs &a, hipStream_t s)
{
    const unsigned n = a.rows * a.cols;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_upsample_copy, grid_for(n / 4 ? n / 4 : 1), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_upsample_filter, grid_for(n), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_rij(const LevelArgs &a, hipStream_t s) { PDF_LAUNCH(k_rij, a.rows * a.cols, a, s); }
hipError_t launch_image_gradients(const LevelArgs &a, hipStream_t s) { PDF_LAUNCH(k_image_gradients, a.rows * a.cols, a, s); }
hipError_t launch_warping(const LevelArgs &a, hipStream_t s) { PDF_LAUNCH(k_warping, a.rows * a.cols, a, s); }
hipError_t launch_mu_steps(const LevelArgs &a, hipStream_t s) { PDF_LAUNCH(k_mu_steps, a.rows * a.cols, a, s); }
hipError_t launch_gradient(const LevelArgs &a, hipStream_t s) { PDF_LAUNCH(k_gradient, a.rows * a.cols, a, s); }
hipError_t launch_dual(const LevelArgs &a, hipStream_t s) { PDF_LAUNCH(k_dual, a.rows * a.cols, a, s); }
hipError_t launch_divergence(const LevelArgs &a, hipStream_t s) { PDF_LAUNCH(k_divergence, a.rows * a.cols, a, s); }
hipError_t launch_primal(const LevelArgs &a, hipStream_t s) { PDF_LAUNCH(k_primal, a.rows * a.cols, a, s); }
hipError_t launch_filter(const LevelArgs &a, hipStream_t s)
{
    const unsigned n = a.rows * a.cols;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_saturate, grid_for(n), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_filter, grid_for(n), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_motion_field(const LevelArgs &a, hipStream_t s) { PDF_LAUNCH(k_motion_field, a.rows * a.cols, a, s); }

// ------------------------------------------------------------------ plain copy
// A streaming 16-byte copy of known size: the calibration point for HBM byte counters
// (every byte read once, written once).
__global__ void __launch_bounds__(256) k_stream_copy(const float4 *src, float4 *dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

hipError_t launch_stream_copy(const void *src, void *dst, size_t bytes, hipStream_t s)
{
    const size_t n = bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks ? blocks : 1), dim3(256), 0, s, (const float4 *)src, (float4 *)dst, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ exact short sequences
// Every float whose bit pattern lies in [lo, hi] goes through one of the short sequences of
// pdflow_pixel.hpp and through the compiler's IEEE operation; out[0] counts the inputs whose
// results differ in any bit, out[1] keeps the smallest such bit pattern (0xffffffff if none).
__global__ void __launch_bounds__(256) k_check_exact(int which, unsigned lo, unsigned hi,
                                                     unsigned long long *count, unsigned *first)
{
    unsigned long long bad = 0;
    unsigned first_bad = 0xffffffffu;
    const unsigned long long n = (unsigned long long)hi - lo + 1;
    for (unsigned long long i = (unsigned long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * 256) {
        const unsigned bits = lo + (unsigned)i;
        const float x = __uint_as_float(bits);
        float fast, ieee;
        if (which == 0) {
            fast = sqrt_rn_ge1(x);
            ieee = sqrtf(x);
        } else if (which == 1) {
            fast = rcp_rn_ge1(x);
            ieee = 1.f / x;
        } else if (which == 2) {
            fast = inv_sqrt_rn_ge1(x);
            ieee = 1.f / sqrtf(x);
        } else if (which == 3) {
            fast = div_rn_half_or_one(1.f, x);
            ieee = 1.f / x;
        } else {
            fast = div_rn_half_or_one(0.5f, x);
            ieee = 0.5f / x;
        }
        if (__float_as_uint(fast) != __float_as_uint(ieee)) {
            bad++;
            first_bad = min(first_bad, bits);
        }
    }
    if (bad) {
        atomicAdd(count, bad);
        atomicMin(first, first_bad);
    }
}

hipError_t check_exact(int which, unsigned lo, unsigned hi, unsigned long long *count, unsigned *first)
{
    hipLaunchKernelGGL(k_check_exact, dim3(256 * 32), dim3(256), 0, nullptr, which, lo, hi, count, first);
    return hipGetLastError();
}

}  // namespace pdf
