// Per-pixel arithmetic shared by the gfx950 kernels (device-only, header-inlined).
// Each function is one piece of a reference device body, evaluated in the same order
// (FP32, no contraction): citations are to pdflow_cudalib.cu.
#pragma once

#include <hip/hip_runtime.h>

namespace pdf {

__device__ __forceinline__ float rhyp(float a, float b) { return 1.f / sqrtf(a * a + b * b); }

__device__ __forceinline__ float sigma_pd_of(float mu_uv, float ddu, float ddv)
{
    return 1.f / (mu_uv * (1.f + fabsf(ddu) + fabsf(ddv)) + 1e-10f);   // .cu:791
}
__device__ __forceinline__ float sigma_p_of(float lam, float r)
{
    return 0.5f / (lam * r + 1e-10f);                                    // .cu:792-795
}
__device__ __forceinline__ float tau_uv_of(float mu_uv, float dd, float lam, float acu_r)
{
    return 1.f / (mu_uv * fabsf(dd) + lam * acu_r + 1e-10f);             // .cu:805-806
}
__device__ __forceinline__ float tau_w_of(float mu_uv, float lam, float acu_r)
{
    return 1.f / (mu_uv + lam * acu_r + 1e-10f);                         // .cu:807
}

__device__ __forceinline__ void project(float &a, float &b)
{
    // unit-disc projection (.cu:849-859)
    const float m = rhyp(a, b);
    if (m < 1.f) { a = a * m; b = b * m; }
}

__device__ __forceinline__ float clamp_pd(float pd)
{
    // .cu:837-846
    if (fabsf(pd) > 1.f) return (pd > 1.f) ? 1.f : -1.f;
    return pd;
}

__device__ __forceinline__ float div_axis(float r0, float p0, float rm, float pm, bool first, bool last)
{
    // one axis of .cu:941-979
    if (first) return r0 * p0;
    if (last) return -rm * pm;
    return r0 * p0 - rm * pm;
}

struct Primal {
    float du, dv, dw;
};

__device__ __forceinline__ Primal primal_update(float du, float dv, float dw, float tu, float tv,
                                                float tw, float muv, float pd, float ddu,
                                                float ddv, float divu, float divv, float divw,
                                                float dct, float dcu, float dcv, float li,
                                                float ld)
{
    // .cu:896-923
    du += -tu * (muv * ddu * pd - li * divu);
    dv += -tv * (muv * ddv * pd - li * divv);
    dw += -tw * (-muv * pd - ld * divw);
    const float of = dct + dcu * du + dcv * dv;
    const float th = tu * dcu * dcu + tv * dcv * dcv;
    if (of < -th) {
        du += tu * dcu;
        dv += tv * dcv;
    } else if (of > th) {
        du -= tu * dcu;
        dv -= tv * dcv;
    } else {
        const float den = tu * dcu * dcu + tv * dcv * dcv + 1e-10f;
        du -= tu * dcu * of / den;
        dv -= tv * dcv * of / den;
    }
    return {du, dv, dw};
}


}  // namespace pdf
