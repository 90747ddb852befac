// Per-pixel arithmetic shared by the gfx950 kernels (device-only, header-inlined).
// Each function is one piece of a reference device body, evaluated in the same order
// (FP32, no contraction): citations are to pdflow_cudalib.cu.
#pragma once

#include <hip/hip_runtime.h>

#include "pdflow_internal.hpp"

namespace pdf {

__device__ __forceinline__ float rhyp(float a, float b) { return 1.f / sqrtf(a * a + b * b); }

// ------------------------------------------------------------------ exact short sequences
// IEEE FP32 sqrt and division on gfx950 are v_sqrt_f32 / v_rcp_f32 estimates made exact by FMA
// steps, wrapped in range handling: a 2^32 pre-scale and class test for sqrt, v_div_scale /
// v_div_fmas / v_div_fixup for division.  On the ranges below the range handling is the
// identity (no operand is tiny, huge, zero, denormal or non-finite).  The square root keeps the
// compiler's FMA steps; for reciprocals of these ranges the v_rcp_f32 estimate after one Newton
// step is already the correctly rounded quotient, so the residual corrections are dropped.
// tests/test_gpu_exact.py sweeps every float of each range against sqrtf and "/" and finds the
// same bits everywhere.

// sqrtf(x) for 1 <= x <= +inf: the estimate s, then the neighbour of s whose residual
// x - s'*s shows that s was one step off (one below, one above).
__device__ __forceinline__ float sqrt_rn_ge1(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float below = __int_as_float(__float_as_int(s) - 1);
    const float above = __int_as_float(__float_as_int(s) + 1);
    const float r_below = __builtin_fmaf(-below, s, x);
    const float r_above = __builtin_fmaf(-above, s, x);
    const float t = r_below <= 0.f ? below : s;
    return r_above > 0.f ? above : t;
}

// 1 / d, correctly rounded for 2^-94 <= d <= 2^124: the v_rcp_f32 estimate refined by one
// Newton step
__device__ __forceinline__ float rcp_rn(float d)
{
    const float r0 = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(__builtin_fmaf(-d, r0, 1.f), r0, r0);
}

// 1.f / d for 1 <= d < 2^64
__device__ __forceinline__ float rcp_rn_ge1(float d) { return rcp_rn(d); }

// n / d for n = 0.5 or 1 and 2^-94 <= d <= 2^124: n times the correctly rounded 1/d (scaling by
// 0.5 is exact in this range)
__device__ __forceinline__ float div_rn_half_or_one(float n, float d) { return n * rcp_rn(d); }

// n / d for n = 0.5 or 1 and any d: the short sequence in its range, the full division (a branch
// no lane takes for the step sizes' positive, guarded denominators) elsewhere
__device__ __forceinline__ float div_half_or_one(float n, float d)
{
    if (__builtin_expect(d >= 0x1p-94f && d <= 0x1p124f, 1)) return div_rn_half_or_one(n, d);
    return n / d;
}

// 1.f / sqrtf(x) for x >= 1 (1/+inf = 0): the reciprocal of rcp_rn with its residual taken on
// min(s, FLT_MAX).  For finite x, s = sqrtf(x) < 2^64 is unchanged by the min; for x = +inf the
// estimate of 1/s is +0 and the Newton step on FLT_MAX leaves it +0, where s itself would give
// inf * 0 = NaN.  s >= 1, so the min is taken on the bit patterns (one integer min, which needs
// no NaN canonicalisation).  No branch and no select.
__device__ __forceinline__ float inv_sqrt_rn_ge1(float x)
{
    const float s = sqrt_rn_ge1(x);
    const float r0 = __builtin_amdgcn_rcpf(s);
    const float sc = __uint_as_float(min(__float_as_uint(s), 0x7f7fffffu));
    return __builtin_fmaf(__builtin_fmaf(-sc, r0, 1.f), r0, r0);
}


__device__ __forceinline__ float sigma_pd_of(float mu_uv, float ddu, float ddv)
{
    return div_half_or_one(1.f, mu_uv * (1.f + fabsf(ddu) + fabsf(ddv)) + 1e-10f);   // .cu:791
}
__device__ __forceinline__ float sigma_p_of(float lam, float r)
{
    return div_half_or_one(0.5f, lam * r + 1e-10f);                      // .cu:792-795
}
__device__ __forceinline__ float tau_uv_of(float mu_uv, float dd, float lam, float acu_r)
{
    return div_half_or_one(1.f, mu_uv * fabsf(dd) + lam * acu_r + 1e-10f);   // .cu:805-806
}
__device__ __forceinline__ float tau_w_of(float mu_uv, float lam, float acu_r)
{
    return div_half_or_one(1.f, mu_uv + lam * acu_r + 1e-10f);           // .cu:807
}

__device__ __forceinline__ void project(float &a, float &b)
{
    // unit-disc projection (.cu:849-859): m = 1/sqrt(a*a+b*b), scale when m < 1.  With
    // correctly rounded sqrt and division, x = a*a+b*b <= 1 gives sqrt(x) <= 1 and so m >= 1,
    // and a NaN x gives a NaN m: neither scales.  Here x is raised to 1 first (fmaxf returns 1
    // for a NaN x), which gives m = 1 exactly in those cases, and the reference's m for every
    // x > 1, +inf included.  So the product by m can be unconditional: a * 1 is a, -0 and
    // infinities included (a NaN stays a NaN).
    const float x = a * a + b * b;
    const float m = inv_sqrt_rn_ge1(fmaxf(x, 1.f));
    a = a * m;
    b = b * m;
}

// project() for a pair that seldom leaves the unit disc (the depth-flow duals pwu, pwv: about 1%
// of the pixels): a wave in which no lane has a*a+b*b > 1 skips it, since project() changes
// nothing then (m >= 1 for x <= 1, and a NaN x is raised to 1).  The test is wave-uniform.
__device__ __forceinline__ void project_if_any(float &a, float &b)
{
    const float x = a * a + b * b;
    if (__builtin_amdgcn_ballot_w64(x > 1.f)) project(a, b);
}

__device__ __forceinline__ float clamp_pd(float pd)
{
    // .cu:837-846: beyond the unit interval the reference returns 1 for pd > 1 and -1 for
    // pd < -1, i.e. 1 with pd's sign (|pd| > 1 excludes NaN)
    return fabsf(pd) > 1.f ? __builtin_copysignf(1.f, pd) : pd;
}

__device__ __forceinline__ float div_axis(float r0, float p0, float rm, float pm, bool first, bool last)
{
    // one axis of .cu:941-979, as selects (-rm * pm is the exact negation of rm * pm)
    const float f = r0 * p0, g = rm * pm;
    const float mid = f - g;
    return first ? f : (last ? -g : mid);
}

// div_axis without the first-element case: the caller passes rm = pm = +0 there, and f - (+0)
// equals f bit for bit (for +0, -0 and every other f)
__device__ __forceinline__ float div_axis_last(float r0, float p0, float rm, float pm, bool last)
{
    const float f = r0 * p0, g = rm * pm;
    const float mid = f - g;
    return last ? -g : mid;
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


__device__ __forceinline__ float wdiff(float rn, float rp, float fn, float f0, float fp)
{
    // r-weighted central difference (.cu:593-595)
    return (rn * (fn - f0) + rp * (f0 - fp)) / (rn + rp);
}

struct Tap {
    unsigned su, fu, sv, fv;
    float iu, iv;
};

__device__ __forceinline__ Tap make_tap(unsigned rows_i, unsigned cols_i, float iu, float iv)
{
    // clamp (.cu:656-659) and the four integer neighbours (.cu:661-664).  A NaN coordinate
    // is taken as 0 so that no value can produce an index outside the plane (the
    // reference leaves that case undefined); finite and infinite inputs are unaffected.
    if (!(iu == iu)) iu = 0.f;
    if (!(iv == iv)) iv = 0.f;
    if (iu < 0.f) iu = 0.f;
    else if (iu > (float)cols_i - 1.f) iu = (float)cols_i - 1.f;
    if (iv < 0.f) iv = 0.f;
    else if (iv > (float)rows_i - 1.f) iv = (float)rows_i - 1.f;
    Tap t;
    t.su = (unsigned)ceilf(iu);
    t.fu = (unsigned)floorf(iu);
    t.sv = (unsigned)ceilf(iv);
    t.fv = (unsigned)floorf(iv);
    t.iu = iu;
    t.iv = iv;
    return t;
}

// The four neighbours (fv, fu), (sv, fu), (fv, su), (sv, su) of a clamped tap in a plane in global
// memory, loaded together.  Every case of interpolatePixel and interpolatePixelDepth reads only
// these: at an integer coordinate the rounded coordinate is that integer, and otherwise rintf
// rounds it to fu or su (fv or sv).
struct Corners {
    float ff, sf, fs, ss;   // (v, u) = (fv, fu), (sv, fu), (fv, su), (sv, su)
};

__device__ __forceinline__ Corners corners(const float *plane, unsigned rows_i, const Tap &t)
{
    const auto mat = gptr(plane);
    return {mat[t.fv + rows_i * t.fu], mat[t.sv + rows_i * t.fu], mat[t.fv + rows_i * t.su],
            mat[t.sv + rows_i * t.su]};
}

__device__ __forceinline__ float bilinear(const Corners &c, const Tap &t)
{
    // interpolatePixel (.cu:666-681) after clamping.  Every case is evaluated with the
    // reference's operations and one is selected, so that no branch holds a corner load:
    // the corners of all the planes a kernel gathers are then in flight together.
    const bool eu = t.su == t.fu, ev = t.sv == t.fv;
    const float au = (float)t.su - t.iu, bu = t.iu - (float)t.fu;
    const float av = (float)t.sv - t.iv, bv = t.iv - (float)t.fv;
    const float col = av * c.ff + bv * c.sf;       // su == fu
    const float bot = au * c.ff + bu * c.fs;       // sv == fv, and the general case's bottom row
    const float top = au * c.sf + bu * c.ss;
    const float both = av * bot + bv * top;
    return eu ? (ev ? c.ff : col) : (ev ? bot : both);
}

__device__ __forceinline__ float bilinear(const float *plane, unsigned rows_i, const Tap &t)
{
    return bilinear(corners(plane, rows_i, t), t);
}

// An empty statement that takes the four corners as operands and clobbers memory: their loads
// stay where they were issued (not sunk into the branches that a later select may become),
// and no load issued before it moves below it.  Placed after all of a kernel's gathers, it
// lets them all be in flight at once.
__device__ __forceinline__ void issued(const Corners &c)
{
    asm volatile("" ::"v"(c.ff), "v"(c.sf), "v"(c.fs), "v"(c.ss) : "memory");
}

__device__ __forceinline__ float bilinear_depth(const Corners &c, const Tap &t)
{
    // interpolatePixelDepth (.cu:696-720): nearest neighbour (half to even) if a tap is 0
    const bool hole = c.ss == 0.f || c.sf == 0.f || c.fs == 0.f || c.ff == 0.f;
    const unsigned ru = (unsigned)rintf(t.iu), rv = (unsigned)rintf(t.iv);
    const float nearest = ru == t.fu ? (rv == t.fv ? c.ff : c.sf) : (rv == t.fv ? c.fs : c.ss);
    const float interp = bilinear(c, t);
    return hole ? nearest : interp;
}

__device__ __forceinline__ float bilinear_depth(const float *plane, unsigned rows_i, const Tap &t)
{
    return bilinear_depth(corners(plane, rows_i, t), t);
}

// Stable sort of 9 (key, weight) pairs by key.  The reference uses a bubble sort with a
// strict ">" (.cu:1457-1473); every stable sort yields the same permutation, so this
// fully unrolled odd-even transposition network (same strict compare, registers only)
// gives the same output.  Unused slots carry key +inf and weight 0 and stay at the end.
__device__ __forceinline__ void sort9(float (&k)[9], float (&w)[9])
{
#pragma unroll
    for (int round = 0; round < 9; round++) {
#pragma unroll
        for (int i = 1 + (round & 1); i < 9; i += 2) {
            if (k[i - 1] > k[i]) {
                const float tk = k[i - 1]; k[i - 1] = k[i]; k[i] = tk;
                const float tw = w[i - 1]; w[i - 1] = w[i]; w[i] = tw;
            }
        }
    }
}

// sort9 of three (key, weight) arrays in lockstep: the same network, each compare-exchange
// done for all three in turn so that their independent instructions interleave
__device__ __forceinline__ void sort9x3(float (&ka)[9], float (&wa)[9], float (&kb)[9],
                                        float (&wb)[9], float (&kc)[9], float (&wc)[9])
{
    auto ce = [](float (&k)[9], float (&w)[9], int i) {
        const bool swap = k[i - 1] > k[i];
        const float lo = swap ? k[i] : k[i - 1], hi = swap ? k[i - 1] : k[i];
        const float wl = swap ? w[i] : w[i - 1], wh = swap ? w[i - 1] : w[i];
        k[i - 1] = lo; k[i] = hi; w[i - 1] = wl; w[i] = wh;
    };
#pragma unroll
    for (int round = 0; round < 9; round++) {
#pragma unroll
        for (int i = 1 + (round & 1); i < 9; i += 2) {
            ce(ka, wa, i);
            ce(kb, wb, i);
            ce(kc, wc, i);
        }
    }
}

__device__ __forceinline__ float wmedian_pick(const float (&fld)[9], const float (&cum)[9], float med)
{
    // .cu:1103-1110: cont = the first index with med <= cum[cont] (at most 8 here).  The
    // values at cont - 1 and cont are carried along while med > cum[i] holds, so no index is
    // formed; cont = 0 returns fld[0].
    bool go = med > cum[0];
    const bool interp = go;
    float cl = cum[0], fl = fld[0], cr = cum[1], fr = fld[1];
#pragma unroll
    for (int i = 1; i < 8; i++) {
        go = go && (med > cum[i]);
        cl = go ? cum[i] : cl;
        fl = go ? fld[i] : fl;
        cr = go ? cum[i + 1] : cr;
        fr = go ? fld[i + 1] : fr;
    }
    const float r = ((cr - med) * fl + (med - cl) * fr) / (cr - cl);
    return interp ? r : fld[0];
}

}  // namespace pdf
