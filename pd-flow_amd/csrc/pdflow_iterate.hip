// The primal-dual iteration (reference loop scene_flow_impair.cpp:138-144: Gradient,
// DualVariables, Divergence, PrimalVariables; bodies pdflow_cudalib.cu:813-1013) as fused
// gfx950 kernels.  The memory-bound hot path of the solver.
//
//   k_pd_fused   one iteration per launch on 32x8 tiles, one pixel per thread.
//   k_pd_tb<K>   K iterations per launch (temporal blocking): a 64x32 region per 1024-thread
//                workgroup keeps all 13 state planes and the 11 constant planes in registers
//                (two pixels per thread), exchanges neighbour values through LDS, and writes
//                back only the inner (64-2K)x(32-2K) tile, whose values after K iterations
//                depend on nothing outside the region.  HBM traffic per pixel-iteration drops
//                from 148 B to about 148*(64*32)/((64-2K)(32-2K))/K B.
//
// Both compute every pixel with the same FP32 operations in the same order as the reference
// kernels, so they agree bit for bit with the unfused stages.  Step sizes are recomputed from
// ri/rj/mu_uv/ddu/ddv (pure functions of constants, identical to what K8 stores).
#include <hip/hip_runtime.h>

#include "pdflow_internal.hpp"
#include "pdflow_pixel.hpp"

namespace pdf {
namespace {

// ------------------------------------------------------------------ one iteration per launch
// The dual step needs (acc+prev) at +1 in v and u; the primal step needs the new duals at -1
// in v and u.  The tile stages (acc+prev) on a 1-pixel ring and recomputes the new TV duals of
// its low-side halo (one row, one column) instead of synchronising with neighbouring tiles.
constexpr int TV = 32, TU = 8;                 // tile rows (v) x cols (u)
constexpr int RV = TV + 2, RU = TU + 2;        // ring region of (acc+prev)
constexpr int HV = TV + 1, HU = TU + 1;        // tile + low halo (ri, rj, new duals)

struct TVDual {
    float puu, puv, pvu, pvv, pwu, pwv;
};

__device__ __forceinline__ void pd_fused_body(const PDArgs &a, unsigned bx, unsigned by)
{
    __shared__ float sU[RU * RV], sV[RU * RV], sW[RU * RV];
    __shared__ float sRi[HU * HV], sRj[HU * HV];
    __shared__ float sP[6][HU * HV];

    const int tid = threadIdx.x;
    const int tv = tid % TV, tu = tid / TV;
    const int v0 = (int)bx * TV, u0 = (int)by * TU;
    const int rows = (int)a.rows, cols = (int)a.cols;
    const float li = a.lambda_i, ld = a.lambda_d;
    const auto du_acc = gptr(a.in[S_DU_ACC]), dv_acc = gptr(a.in[S_DV_ACC]), dw_acc = gptr(a.in[S_DW_ACC]);

    // own pixel: issue every global load before the first barrier
    const int gv = v0 + tv, gu = u0 + tu;
    const bool inside = gv < rows && gu < cols;
    const unsigned g = inside ? (unsigned)gv + (unsigned)gu * a.rows : 0u;
    float ua = 0, va = 0, wa = 0, du0 = 0, dv0 = 0, dw0 = 0, pd_in = 0, mu_uv = 0, ddu = 0,
          ddv = 0, ddt = 0, dct = 0, dcu = 0, dcv = 0;
    float p_in[6] = {0, 0, 0, 0, 0, 0};
    if (inside) {
        ua = du_acc[g]; va = dv_acc[g]; wa = dw_acc[g];
        du0 = gptr(a.in[S_DU])[g]; dv0 = gptr(a.in[S_DV])[g]; dw0 = gptr(a.in[S_DW])[g]; pd_in = gptr(a.in[S_PD])[g];
#pragma unroll
        for (int k = 0; k < 6; k++) p_in[k] = gptr(a.in[S_PUU + k])[g];
        mu_uv = gptr(a.mu_uv)[g]; ddu = gptr(a.ddu)[g]; ddv = gptr(a.ddv)[g]; ddt = gptr(a.ddt)[g];
        dct = gptr(a.dct)[g]; dcu = gptr(a.dcu)[g]; dcv = gptr(a.dcv)[g];
    }

    // stage (acc + prev) on the ring region [v0-1, v0+TV] x [u0-1, u0+TU]
    for (int e = tid; e < RU * RV; e += TV * TU) {
        const int ev = e % RV - 1, eu = e / RV - 1;
        const int qv = v0 + ev, qu = u0 + eu;
        float su = 0.f, sv = 0.f, sw = 0.f;
        if (qv >= 0 && qv < rows && qu >= 0 && qu < cols) {
            const unsigned q = (unsigned)qv + (unsigned)qu * a.rows;
            su = du_acc[q] + gptr(a.du_prev)[q];
            sv = dv_acc[q] + gptr(a.dv_prev)[q];
            sw = dw_acc[q];
        }
        sU[e] = su; sV[e] = sv; sW[e] = sw;
    }
    // stage ri, rj on tile + low halo [v0-1, v0+TV-1] x [u0-1, u0+TU-1]
    for (int e = tid; e < HU * HV; e += TV * TU) {
        const int ev = e % HV - 1, eu = e / HV - 1;
        const int qv = v0 + ev, qu = u0 + eu;
        float ri = 0.f, rj = 0.f;
        if (qv >= 0 && qv < rows && qu >= 0 && qu < cols) {
            const unsigned q = (unsigned)qv + (unsigned)qu * a.rows;
            ri = gptr(a.ri)[q];
            rj = gptr(a.rj)[q];
        }
        sRi[e] = ri; sRj[e] = rj;
    }
    // the low-halo pixel this thread recomputes (row v0-1 or column u0-1), if any
    int hev = 0, heu = 0;
    bool halo = false;
    float hp[6] = {0, 0, 0, 0, 0, 0};
    if (tid < TU + TV) {
        if (tid < TU) { hev = -1; heu = tid; }
        else { hev = tid - TU; heu = -1; }
        const int hv = v0 + hev, hu = u0 + heu;
        halo = hv >= 0 && hu >= 0 && hv < rows && hu < cols;
        if (halo) {
            const unsigned hg = (unsigned)hv + (unsigned)hu * a.rows;
#pragma unroll
            for (int k = 0; k < 6; k++) hp[k] = gptr(a.in[S_PUU + k])[hg];
        }
    }
    __syncthreads();

    // dual step of the TV duals at a pixel of the tile or its low halo (local ev, eu)
    auto tv_dual = [&](int ev, int eu, int qv, int qu, const float *pin) -> TVDual {
        const int r = (eu + 1) * RV + (ev + 1);          // ring index
        const int h = (eu + 1) * HV + (ev + 1);          // halo-grid index
        const float ri = sRi[h], rj = sRj[h];
        const float u_c = sU[r], v_c = sV[r], w_c = sW[r];
        float g1u = 0.f, g1v = 0.f, g1w = 0.f, g2u = 0.f, g2v = 0.f, g2w = 0.f;
        if (qu != cols - 1) {
            g1u = ri * (sU[r + RV] - u_c);
            g1v = ri * (sV[r + RV] - v_c);
            g1w = ri * (sW[r + RV] - w_c);
        }
        if (qv != rows - 1) {
            g2u = rj * (sU[r + 1] - u_c);
            g2v = rj * (sV[r + 1] - v_c);
            g2w = rj * (sW[r + 1] - w_c);
        }
        const float sx = sigma_p_of(li, ri), sy = sigma_p_of(li, rj);
        const float swx = div_half_or_one(0.5f, ri * ld + 1e-10f), swy = div_half_or_one(0.5f, rj * ld + 1e-10f);
        TVDual p;
        p.puu = pin[0] + sx * li * g1u;
        p.puv = pin[1] + sy * li * g2u;
        p.pvu = pin[2] + sx * li * g1v;
        p.pvv = pin[3] + sy * li * g2v;
        p.pwu = pin[4] + swx * ld * g1w;
        p.pwv = pin[5] + swy * ld * g2w;
        project(p.puu, p.puv);
        project(p.pvu, p.pvv);
        project_if_any(p.pwu, p.pwv);
        return p;
    };
    auto store_dual = [&](int h, const TVDual &p) {
        sP[0][h] = p.puu; sP[1][h] = p.puv; sP[2][h] = p.pvu;
        sP[3][h] = p.pvv; sP[4][h] = p.pwu; sP[5][h] = p.pwv;
    };

    const int h = (tu + 1) * HV + (tv + 1);
    float pd = 0.f;
    if (inside) {
        const TVDual p = tv_dual(tv, tu, gv, gu, p_in);
        store_dual(h, p);
        gptr(a.out[S_PUU])[g] = p.puu; gptr(a.out[S_PUV])[g] = p.puv; gptr(a.out[S_PVU])[g] = p.pvu;
        gptr(a.out[S_PVV])[g] = p.pvv; gptr(a.out[S_PWU])[g] = p.pwu; gptr(a.out[S_PWV])[g] = p.pwv;
        pd = pd_in + sigma_pd_of(mu_uv, ddu, ddv) * mu_uv * (-wa + ddt + ddu * ua + ddv * va);
        pd = clamp_pd(pd);
        gptr(a.out[S_PD])[g] = pd;
    }
    if (halo) store_dual((heu + 1) * HV + (hev + 1), tv_dual(hev, heu, v0 + hev, u0 + heu, hp));
    __syncthreads();

    if (inside) {
        const float ri = sRi[h], rj = sRj[h];
        const float riu = sRi[h - HV], rjv = sRj[h - 1];   // (v, u-1) and (v-1, u)
        const bool uf = gu == 0, ul = gu == cols - 1, vf = gv == 0, vl = gv == rows - 1;
        float divu = div_axis(ri, sP[0][h], riu, sP[0][h - HV], uf, ul);
        float divv = div_axis(ri, sP[2][h], riu, sP[2][h - HV], uf, ul);
        float divw = div_axis(ri, sP[4][h], riu, sP[4][h - HV], uf, ul);
        divu += div_axis(rj, sP[1][h], rjv, sP[1][h - 1], vf, vl);
        divv += div_axis(rj, sP[3][h], rjv, sP[3][h - 1], vf, vl);
        divw += div_axis(rj, sP[5][h], rjv, sP[5][h - 1], vf, vl);
        float acu_r = ri + rj;
        if (gu > 0) acu_r += riu;
        if (gv > 0) acu_r += rjv;
        const Primal p = primal_update(du0, dv0, dw0, tau_uv_of(mu_uv, ddu, li, acu_r),
                                       tau_uv_of(mu_uv, ddv, li, acu_r), tau_w_of(mu_uv, ld, acu_r),
                                       mu_uv, pd, ddu, ddv, divu, divv, divw, dct, dcu, dcv, li, ld);
        gptr(a.out[S_DU_ACC])[g] = 2.f * p.du - du0;
        gptr(a.out[S_DV_ACC])[g] = 2.f * p.dv - dv0;
        gptr(a.out[S_DW_ACC])[g] = 2.f * p.dw - dw0;
        gptr(a.out[S_DU])[g] = p.du;
        gptr(a.out[S_DV])[g] = p.dv;
        gptr(a.out[S_DW])[g] = p.dw;
    }
}

// ------------------------------------------------------------------ register-resident pixels
// A pixel whose whole iteration state lives in registers across iterations (the
// temporally blocked and whole-level kernels).  Neighbour values travel through LDS.
struct Px {
    // state
    float du, dv, dw, ua, va, wa, pd, p[6];
    // constants
    float up, vp, mu, ddu, ddv, ddt, dct, dcu, dcv, ri, rj;
    // per-launch derived constants (pure functions of the constants)
    float sxl, syl, swxl, swyl, spdm, tu, tv, tw;
    int gv, gu;
    bool on;
};

// Loads are unconditional: pixels outside the image load pixel 0's values, which no pixel
// inside the image reads (every image-border term is selected away by gv/gu).  The empty asm
// statement makes the selected index opaque: otherwise the compiler widens the select to 64 bits
// and adds each plane's base with a 64-bit VALU add, where a load can take the base in SGPRs and
// the 32-bit byte offset in one VGPR (one VALU instruction fewer per plane).
__device__ __forceinline__ unsigned px_index(const PDArgs &a, int gv, int gu, bool on)
{
    unsigned g = on ? (unsigned)gv + (unsigned)gu * a.rows : 0u;
    asm("" : "+v"(g));
    return g;
}

// the 13 state planes, ri and rj
// FRESH: the first launch of a level whose set-up wrote only dw of the primal state: du, dv and
// their accumulators start at 0 and dw_acc at dw (.cu:447-454, 556-568), set here instead of
// loaded
template <bool FRESH = false>
__device__ __forceinline__ void load_state(Px &q, const PDArgs &a, int gv, int gu, bool on)
{
    q.gv = gv;
    q.gu = gu;
    q.on = on;
    const unsigned g = px_index(a, gv, gu, on);
    if constexpr (FRESH) {
        q.dw = gat(a.in[S_DW], g);
        q.du = 0.f; q.dv = 0.f; q.ua = 0.f; q.va = 0.f; q.wa = q.dw;
    } else {
        q.du = gat(a.in[S_DU], g); q.dv = gat(a.in[S_DV], g); q.dw = gat(a.in[S_DW], g);
        q.ua = gat(a.in[S_DU_ACC], g); q.va = gat(a.in[S_DV_ACC], g); q.wa = gat(a.in[S_DW_ACC], g);
    }
    q.pd = gat(a.in[S_PD], g);
#pragma unroll
    for (int k = 0; k < 6; k++) q.p[k] = gat(a.in[S_PUU + k], g);
    q.ri = gat(a.ri, g); q.rj = gat(a.rj, g);
}

// the other 9 per-level constant planes
constexpr int kConstPlanes = 9;
__device__ __forceinline__ const float *const_plane(const PDArgs &a, int k)
{
    const float *const planes[kConstPlanes] = {a.du_prev, a.dv_prev, a.mu_uv, a.ddu, a.ddv,
                                               a.ddt, a.dct, a.dcu, a.dcv};
    return planes[k];
}
__device__ __forceinline__ void set_consts(Px &q, const float (&c)[kConstPlanes])
{
    q.up = c[0]; q.vp = c[1]; q.mu = c[2]; q.ddu = c[3]; q.ddv = c[4];
    q.ddt = c[5]; q.dct = c[6]; q.dcu = c[7]; q.dcv = c[8];
}

template <bool FRESH = false>
__device__ __forceinline__ void load_px(Px &q, const PDArgs &a, int gv, int gu, bool on)
{
    load_state<FRESH>(q, a, gv, gu, on);
    const unsigned g = px_index(a, gv, gu, on);
    float c[kConstPlanes];
#pragma unroll
    for (int k = 0; k < kConstPlanes; k++) c[k] = gat(const_plane(a, k), g);
    set_consts(q, c);
}

__device__ __forceinline__ void store_px(const Px &q, const PDArgs &a)
{
    const unsigned g = (unsigned)q.gv + (unsigned)q.gu * a.rows;
    gat(a.out[S_DU], g) = q.du; gat(a.out[S_DV], g) = q.dv; gat(a.out[S_DW], g) = q.dw;
    gat(a.out[S_DU_ACC], g) = q.ua; gat(a.out[S_DV_ACC], g) = q.va; gat(a.out[S_DW_ACC], g) = q.wa;
    gat(a.out[S_PD], g) = q.pd;
#pragma unroll
    for (int k = 0; k < 6; k++) gat(a.out[S_PUU + k], g) = q.p[k];
}

// riu = ri at (v, u-1), rjv = rj at (v-1, u); only read when those pixels exist
__device__ __forceinline__ void derive_px(Px &q, float riu, float rjv, float li, float ld)
{
    // the first products of the reference expressions (.cu:822-834) are per-launch constants
    q.sxl = sigma_p_of(li, q.ri) * li;
    q.syl = sigma_p_of(li, q.rj) * li;
    q.swxl = div_half_or_one(0.5f, q.ri * ld + 1e-10f) * ld;
    q.swyl = div_half_or_one(0.5f, q.rj * ld + 1e-10f) * ld;
    q.spdm = sigma_pd_of(q.mu, q.ddu, q.ddv) * q.mu;
    float acu_r = q.ri + q.rj;                                   // .cu:801-803
    if (q.gu > 0) acu_r += riu;
    if (q.gv > 0) acu_r += rjv;
    q.tu = tau_uv_of(q.mu, q.ddu, li, acu_r);
    q.tv = tau_uv_of(q.mu, q.ddv, li, acu_r);
    q.tw = tau_w_of(q.mu, ld, acu_r);
}

// dual step (.cu:813-887 with the gradient of .cu:982-1013): n* are (acc+prev) at (v, u+1),
// m* at (v+1, u).  In the last column (row) the reference's gradient is +0; the caller passes the
// pixel's own values there (the same bits as u_c, v_c, w_c), whose difference is +0 and whose
// product with the positive weight is +0 as well, so no border case is needed here (for finite
// state values).
__device__ __forceinline__ void dual_px(Px &q, float nU, float nV, float nW, float mU, float mV, float mW)
{
    const float u_c = q.ua + q.up, v_c = q.va + q.vp, w_c = q.wa;
    const float g1u = q.ri * (nU - u_c), g1v = q.ri * (nV - v_c), g1w = q.ri * (nW - w_c);
    const float g2u = q.rj * (mU - u_c), g2v = q.rj * (mV - v_c), g2w = q.rj * (mW - w_c);
    q.p[0] = q.p[0] + q.sxl * g1u;
    q.p[1] = q.p[1] + q.syl * g2u;
    q.p[2] = q.p[2] + q.sxl * g1v;
    q.p[3] = q.p[3] + q.syl * g2v;
    q.p[4] = q.p[4] + q.swxl * g1w;
    q.p[5] = q.p[5] + q.swyl * g2w;
    project(q.p[0], q.p[1]);
    project(q.p[2], q.p[3]);
    project_if_any(q.p[4], q.p[5]);
    q.pd = clamp_pd(q.pd + q.spdm * (-q.wa + q.ddt + q.ddu * q.ua + q.ddv * q.va));
}

// primal step (.cu:889-933 with the divergence of .cu:935-980): pu* are the new duals
// puu, pvu, pwu at (v, u-1) with riu = ri there; pv* are puv, pvv, pwv at (v-1, u) with
// rjv = rj there.  In the first column (row) the caller passes +0 for the weight and the duals
// (an LDS slot that holds +0): the product is +0 and f - (+0) is f bit for bit, the
// reference's first-column term.
// EDGE = false: the caller knows that no pixel of the wave lies in the image's last column or
// row, where the divergence takes its one-sided form, so that case is left out
template <bool EDGE = true>
__device__ __forceinline__ void primal_px(Px &q, int rows, int cols, float riu, float puu,
                                          float pvu, float pwu, float rjv, float puv, float pvv,
                                          float pwv, float li, float ld)
{
    const bool ul = EDGE && q.gu == cols - 1, vl = EDGE && q.gv == rows - 1;
    float divu = div_axis_last(q.ri, q.p[0], riu, puu, ul);
    float divv = div_axis_last(q.ri, q.p[2], riu, pvu, ul);
    float divw = div_axis_last(q.ri, q.p[4], riu, pwu, ul);
    divu += div_axis_last(q.rj, q.p[1], rjv, puv, vl);
    divv += div_axis_last(q.rj, q.p[3], rjv, pvv, vl);
    divw += div_axis_last(q.rj, q.p[5], rjv, pwv, vl);
    const float du0 = q.du, dv0 = q.dv, dw0 = q.dw;
    const Primal p = primal_update(du0, dv0, dw0, q.tu, q.tv, q.tw, q.mu, q.pd, q.ddu, q.ddv, divu,
                                   divv, divw, q.dct, q.dcu, q.dcv, li, ld);
    q.ua = 2.f * p.du - du0;
    q.va = 2.f * p.dv - dv0;
    q.wa = 2.f * p.dw - dw0;
    q.du = p.du;
    q.dv = p.dv;
    q.dw = p.dw;
}

// ------------------------------------------------------------------ K iterations per launch
// Region shapes (rows RV x columns RU, 1024 threads; lanes of a wave run down RV consecutive
// rows of one or two column pairs):
//   shape 0: 64 x 32, two pixels per thread in adjacent columns, a wave = one column pair
//            (one workgroup per CU: 88 KB LDS, ~120 VGPRs)
//   shape 1: 32 x 32, one pixel per thread, a wave = the two columns of one pair (two
//            workgroups per CU, 45 KB LDS each, ~62 VGPRs, so 8 waves per SIMD hide the
//            latency of the arithmetic and one workgroup's loads overlap the other's work)
//
// Work that no output needs is skipped a whole wave at a time.  With d a pixel's distance to
// the region's nearest edge (in rows or columns), iteration `it` (0-based) needs the dual step
// at d >= it and the primal step at d >= it + 1: the dual at (v, u) reads ubar at (v, u+1) and
// (v+1, u), the primal reads the new duals at (v, u-1) and (v-1, u), so each iteration widens
// the band of wrong values by one pixel on every side, and the stored tile is d >= K.  The
// column pairs are dealt to waves from both ends inwards (wave w holds pairs w and
// RU/2 - 1 - w when a wave holds two), so the waves at the left and right edges hold only
// small d and drop out after their first iterations.
constexpr int BT = 1024;               // threads of the whole-level kernel

// threads of a temporally blocked workgroup, and pixels per thread of a region shape
constexpr int kTBThreads = 1024;
template <int RV, int RU>
constexpr int tb_px() { return RV * RU / kTBThreads; }
// waves per SIMD the register budget must allow: two 16-wave workgroups per CU with one pixel
// per thread (at most 64 VGPRs), one workgroup with two
template <int RV, int RU>
constexpr int tb_waves_per_simd() { return tb_px<RV, RU>() == 1 ? 8 : 4; }

// column pair (columns 2q, 2q+1 of the region) of wave w: pairs are dealt from both ends
// inwards (0, NP-1, 1, NP-2, ...), one pair per wave
template <int NP>
__device__ __forceinline__ int tb_pair(int w)
{
    return (w & 1) ? NP - 1 - (w >> 1) : (w >> 1);
}

// largest distance to the region edge over the pixels of a wave that holds column pair q and
// spans all RV rows
template <int RV, int RU>
__device__ __forceinline__ int tb_pair_reach(int q)
{
    const int c = 2 * q;
    return min(max(min(c, RU - 1 - c), min(c + 1, RU - 2 - c)), (RV - 1) / 2);
}

// applies f(pixel, j) to a thread's two pixels (named variables, so they stay in registers)
#define PD_BOTH(f) do { f(x0, 0); f(x1, 1); } while (0)
// applies f(pixel, j) to a thread's pixels under a region map m: column 2 * m.slot + j
#define PD_EACH(f) do { if constexpr (PX == 2) { f(x0, 0); f(x1, 1); } else { f(x0, m.j); } } while (0)

// neighbour exchange of one region (11 planes)
template <int RV, int RU>
struct TBLds {
    static constexpr int kZero = RU * RV;      // a slot of ri, rj and the duals that holds +0
    float u[RU * RV], v[RU * RV], w[RU * RV];
    float p[6][RU * RV + 1];
    float ri[RU * RV + 1], rj[RU * RV + 1];
};

// where a thread sits in its region: a wave holds one column pair (slot) down all RV rows;
// with two pixels per thread (RV = 64) a thread holds both columns of its row, with one
// (RV = 32) lanes 0-31 hold column 2 * slot and lanes 32-63 column 2 * slot + 1 (j)
template <int RV, int RU>
struct TBMap {
    int lv, slot, reach, j;
    __device__ __forceinline__ TBMap()
    {
        constexpr int PX = tb_px<RV, RU>(), NP = RU / 2;
        static_assert(PX * 64 == 2 * RV && RV * RU / PX / 64 == NP, "one column pair per wave");
        const int lane = threadIdx.x & 63;
        const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
        lv = lane % RV;
        j = PX == 1 ? lane / RV : 0;
        slot = tb_pair<NP>(wave);
        reach = tb_pair_reach<RV, RU>(slot);                     // wave-uniform
    }
};

// The K iterations of one region whose origin (rv0, ru0) is K rows and columns before its
// output tile: both pixels are loaded and their ri, rj are in s.ri / s.rj behind a barrier.
// Derives the step sizes, iterates, and writes back the inner tile.
template <int K, int RV, int RU>
__device__ __forceinline__ void tb_iterate(const PDArgs &a, TBLds<RV, RU> &s, const TBMap<RV, RU> &m,
                                           Px &x0, Px &x1)
{
    constexpr int TVk = RV - 2 * K, TUk = RU - 2 * K, PX = tb_px<RV, RU>();
    const int rows = (int)a.rows, cols = (int)a.cols;
    const float li = a.lambda_i, ld = a.lambda_d;
    const int lv = m.lv, slot = m.slot, reach = m.reach;
    // neighbours outside the region only feed pixels of the discarded band, so their
    // indices are clamped into the region
    auto derive = [&](Px &q, int j) __attribute__((always_inline)) {
        const int lu = 2 * slot + j, idx = lu * RV + lv;
        derive_px(q, s.ri[lu > 0 ? idx - RV : idx], s.rj[lv > 0 ? idx - 1 : idx], li, ld);
    };
    PD_EACH(derive);

    auto put_bar = [&](Px &q, int j) __attribute__((always_inline)) {
        const int idx = (2 * slot + j) * RV + lv;
        s.u[idx] = q.ua + q.up;
        s.v[idx] = q.va + q.vp;
        s.w[idx] = q.wa;
    };
    // pixels outside the image run the same arithmetic on the values loaded for them; no pixel
    // inside reads their results (the image-border terms are selected away), and without a
    // branch the two pixels' instruction streams interleave
    // neighbour slots: the pixel's own slot stands for (v, u+1) / (v+1, u) in the last column /
    // row of the image (and at the region's edge, where only the discarded band reads it); the
    // +0 slot stands for (v, u-1) / (v-1, u) in the first column / row (and at the region's edge)
    auto dual = [&](Px &q, int j) __attribute__((always_inline)) {
        const int lu = 2 * slot + j, idx = lu * RV + lv;
        const int iu = lu + 1 < RU && q.gu != cols - 1 ? idx + RV : idx;
        const int iv = lv + 1 < RV && q.gv != rows - 1 ? idx + 1 : idx;
        dual_px(q, s.u[iu], s.v[iu], s.w[iu], s.u[iv], s.v[iv], s.w[iv]);
#pragma unroll
        for (int k = 0; k < 6; k++) s.p[k][idx] = q.p[k];
    };
    // most waves hold no pixel of the image's last column or row: they take the primal step
    // without that case's selects (a wave-uniform branch)
    bool edge = false;
    auto find_edge = [&](Px &q, int) __attribute__((always_inline)) {
        edge = edge || __builtin_amdgcn_ballot_w64(q.gu == cols - 1 || q.gv == rows - 1) != 0;
    };
    PD_EACH(find_edge);
    auto primal = [&](Px &q, int j) __attribute__((always_inline)) {
        const int lu = 2 * slot + j, idx = lu * RV + lv;
        const int ju = lu > 0 && q.gu > 0 ? idx - RV : s.kZero;
        const int jv = lv > 0 && q.gv > 0 ? idx - 1 : s.kZero;
        if (edge)
            primal_px<true>(q, rows, cols, s.ri[ju], s.p[0][ju], s.p[2][ju], s.p[4][ju], s.rj[jv],
                            s.p[1][jv], s.p[3][jv], s.p[5][jv], li, ld);
        else
            primal_px<false>(q, rows, cols, s.ri[ju], s.p[0][ju], s.p[2][ju], s.p[4][ju], s.rj[jv],
                             s.p[1][jv], s.p[3][jv], s.p[5][jv], li, ld);
    };
    for (int it = 0; it < K; it++) {
        if (reach >= it) PD_EACH(put_bar);
        __syncthreads();
        if (reach >= it) PD_EACH(dual);
        __syncthreads();
        if (reach > it) PD_EACH(primal);
    }

    // write back the inner tile
    auto store = [&](Px &q, int j) __attribute__((always_inline)) {
        const int lu = 2 * slot + j;
        if (q.on && lv >= K && lv < K + TVk && lu >= K && lu < K + TUk) store_px(q, a);
    };
    PD_EACH(store);
}

template <int K, int RV, int RU, bool FRESH = false>
__device__ __forceinline__ void tb_body(const PDArgs &a, unsigned bx, unsigned by)
{
    constexpr int TVk = RV - 2 * K, TUk = RU - 2 * K, PX = tb_px<RV, RU>();
    static_assert(TVk > 0 && TUk > 0, "region shape");
    __shared__ TBLds<RV, RU> s;
    const TBMap<RV, RU> m;
    const int rv0 = (int)bx * TVk - K, ru0 = (int)by * TUk - K;  // region origin
    const int rows = (int)a.rows, cols = (int)a.cols;

    Px x0, x1;
    auto load = [&](Px &q, int j) __attribute__((always_inline)) {
        const int lu = 2 * m.slot + j;
        const int gv = rv0 + m.lv, gu = ru0 + lu;
        load_px<FRESH>(q, a, gv, gu, gv >= 0 && gv < rows && gu >= 0 && gu < cols);
        s.ri[lu * RV + m.lv] = q.ri;
        s.rj[lu * RV + m.lv] = q.rj;
    };
    PD_EACH(load);
    if (threadIdx.x < 8) {
        // the +0 slot of ri, rj and the six dual planes (the dual steps never write it)
        float *z = threadIdx.x < 6 ? s.p[threadIdx.x] : (threadIdx.x == 6 ? s.ri : s.rj);
        z[s.kZero] = 0.f;
    }
    __syncthreads();
    // most regions lie inside the image: they iterate without the border cases
    tb_iterate<K, RV, RU>(a, s, m, x0, x1);
}

// ------------------------------------------------------------------ whole level per launch
// Levels of at most 2048 pixels (15x20, 30x40) fit one 1024-thread workgroup with two pixels
// per thread in column-major order: all n iterations run in one launch, with no halo since
// the image borders are the only borders.
constexpr int kLevelMaxPixels = 2 * BT;

__device__ __forceinline__ void level_body(const PDArgs &a, int n)
{
    __shared__ float sU[kLevelMaxPixels], sV[kLevelMaxPixels], sW[kLevelMaxPixels];
    // one more slot of ri, rj and the duals, holding +0 (see primal_px)
    constexpr int kZero = kLevelMaxPixels;
    __shared__ float sP[6][kLevelMaxPixels + 1];
    __shared__ float sRi[kLevelMaxPixels + 1], sRj[kLevelMaxPixels + 1];
    const int rows = (int)a.rows, cols = (int)a.cols, P = rows * cols;
    const float li = a.lambda_i, ld = a.lambda_d;

    Px x0, x1;
    auto load = [&](Px &q, int j) __attribute__((always_inline)) {
        const int p = (int)threadIdx.x + j * BT;
        const bool on = p < P;
        load_px(q, a, on ? p % rows : 0, on ? p / rows : 0, on);
        if (on) {
            sRi[p] = q.ri;
            sRj[p] = q.rj;
        }
    };
    PD_BOTH(load);
    if (threadIdx.x < 8) {
        float *z = threadIdx.x < 6 ? sP[threadIdx.x] : (threadIdx.x == 6 ? sRi : sRj);
        z[kZero] = 0.f;
    }
    __syncthreads();
    // a pixel reads (v, u-1) only if u > 0 and (v-1, u) only if v > 0: indices stay in [0, P)
    auto derive = [&](Px &q, int j) __attribute__((always_inline)) {
        const int p = (int)threadIdx.x + j * BT;
        if (q.on) derive_px(q, sRi[q.gu > 0 ? p - rows : p], sRj[q.gv > 0 ? p - 1 : p], li, ld);
    };
    PD_BOTH(derive);
    auto put_bar = [&](Px &q, int j) __attribute__((always_inline)) {
        const int p = (int)threadIdx.x + j * BT;
        if (q.on) {
            sU[p] = q.ua + q.up;
            sV[p] = q.va + q.vp;
            sW[p] = q.wa;
        }
    };
    auto dual = [&](Px &q, int j) __attribute__((always_inline)) {
        const int p = (int)threadIdx.x + j * BT;
        if (!q.on) return;
        const int iu = q.gu != cols - 1 ? p + rows : p;
        const int iv = q.gv != rows - 1 ? p + 1 : p;
        dual_px(q, sU[iu], sV[iu], sW[iu], sU[iv], sV[iv], sW[iv]);
#pragma unroll
        for (int k = 0; k < 6; k++) sP[k][p] = q.p[k];
    };
    auto primal = [&](Px &q, int j) __attribute__((always_inline)) {
        const int p = (int)threadIdx.x + j * BT;
        if (!q.on) return;
        const int ju = q.gu > 0 ? p - rows : kZero;
        const int jv = q.gv > 0 ? p - 1 : kZero;
        primal_px(q, rows, cols, sRi[ju], sP[0][ju], sP[2][ju], sP[4][ju], sRj[jv], sP[1][jv],
                  sP[3][jv], sP[5][jv], li, ld);
    };
    // per iteration: the barrier after put_bar orders the previous primal reads of sP before
    // this dual's writes; the barrier after dual orders this dual's reads of sU/sV/sW before
    // the next put_bar
    for (int it = 0; it < n; it++) {
        PD_BOTH(put_bar);
        __syncthreads();
        PD_BOTH(dual);
        __syncthreads();
        PD_BOTH(primal);
    }
    auto store = [&](Px &q, int) __attribute__((always_inline)) {
        if (q.on) store_px(q, a);
    };
    PD_BOTH(store);
}

// ------------------------------------------------------------------ entry points
__global__ void __launch_bounds__(TV * TU) k_pd_fused(PDArgs a) { pd_fused_body(a, blockIdx.x, blockIdx.y); }
template <int K, int RV, int RU>
__global__ void __launch_bounds__(kTBThreads) __attribute__((amdgpu_waves_per_eu(tb_waves_per_simd<RV, RU>())))
k_pd_tb(PDArgs a)
{
    // tiles numbered v-major, visited in XCD-contiguous runs
    const unsigned t = xcd_swizzle(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    tb_body<K, RV, RU>(a, t % gridDim.x, t / gridDim.x);
}
__global__ void __launch_bounds__(BT) k_pd_level(PDArgs a, int n) { level_body(a, n); }

// batched: every table entry is one pair's PDArgs; tiles of a pair are numbered v-major
__global__ void __launch_bounds__(TV * TU) kb_pd_fused(BatchTable t)
{
    const BatchItem it = batch_item(t);
    const PDArgs &a = static_cast<const PDArgs *>(t.args)[it.item];
    const unsigned tiles_v = (a.rows + TV - 1) / TV;
    pd_fused_body(a, it.block % tiles_v, it.block / tiles_v);
}
template <int K, int RV, int RU, bool FRESH>
__global__ void __launch_bounds__(kTBThreads) __attribute__((amdgpu_waves_per_eu(tb_waves_per_simd<RV, RU>())))
kb_pd_tb(BatchTable t)
{
    const BatchItem it = batch_item(t, xcd_swizzle(blockIdx.x, gridDim.x));
    const PDArgs &a = static_cast<const PDArgs *>(t.args)[it.item];
    const unsigned tiles_v = (a.rows + (RV - 2 * K) - 1) / (RV - 2 * K);
    tb_body<K, RV, RU, FRESH>(a, it.block % tiles_v, it.block / tiles_v);
}
__global__ void __launch_bounds__(BT) kb_pd_level(BatchTable t)
{
    const BatchItem it = batch_item(t);
    const PDArgs &a = static_cast<const PDArgs *>(t.args)[it.item];
    level_body(a, (int)a.iters);
}

}  // namespace

hipError_t launch_pd_fused(const PDArgs &a, hipStream_t s)
{
    if (a.rows == 0 || a.cols == 0) return hipSuccess;
    const dim3 grid((a.rows + TV - 1) / TV, (a.cols + TU - 1) / TU);
    hipLaunchKernelGGL(k_pd_fused, grid, dim3(TV * TU), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pd_level(const PDArgs &a, unsigned n, hipStream_t s)
{
    if (a.rows * a.cols > (unsigned)kLevelMaxPixels) return hipErrorInvalidValue;
    if (a.rows == 0 || a.cols == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pd_level, dim3(1), dim3(BT), 0, s, a, (int)n);
    return hipGetLastError();
}

bool pd_level_supported(unsigned pixels) { return pixels <= (unsigned)kLevelMaxPixels; }

bool pd_tb_supported(int k) { return k == 2 || k == 3 || k == 4 || k == 5 || k == 6 || k == 8 || k == 10 || k == 12; }
bool pd_tb_shape_supported(int shape) { return shape == kTBShape64x32 || shape == kTBShape32x32; }

template <int RV, int RU>
static void tb_region(int k, unsigned rows, unsigned cols, unsigned *tv, unsigned *tu)
{
    *tv = (rows + (RV - 2 * k) - 1) / (RV - 2 * k);
    *tu = (cols + (RU - 2 * k) - 1) / (RU - 2 * k);
}

template <int K, int RV, int RU>
static hipError_t launch_tb(const PDArgs &a, hipStream_t s)
{
    unsigned tv, tu;
    tb_region<RV, RU>(K, a.rows, a.cols, &tv, &tu);
    hipLaunchKernelGGL((k_pd_tb<K, RV, RU>), dim3(tv, tu), dim3(kTBThreads), 0, s, a);
    return hipGetLastError();
}

template <int RV, int RU>
static hipError_t launch_tb_shape(const PDArgs &a, int k, hipStream_t s)
{
    switch (k) {
    case 2: return launch_tb<2, RV, RU>(a, s);
    case 3: return launch_tb<3, RV, RU>(a, s);
    case 4: return launch_tb<4, RV, RU>(a, s);
    case 5: return launch_tb<5, RV, RU>(a, s);
    case 6: return launch_tb<6, RV, RU>(a, s);
    case 8: return launch_tb<8, RV, RU>(a, s);
    case 10: return launch_tb<10, RV, RU>(a, s);
    case 12: return launch_tb<12, RV, RU>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pd_tb(const PDArgs &a, int k, int shape, hipStream_t s)
{
    if (a.rows == 0 || a.cols == 0) return hipSuccess;
    switch (shape) {
    case kTBShape64x32: return launch_tb_shape<64, 32>(a, k, s);
    case kTBShape32x32: return launch_tb_shape<32, 32>(a, k, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pdf

namespace pdf {

unsigned pd_blocks(int k, int shape, unsigned rows, unsigned cols)
{
    if (k <= 0) return 1;                                   // whole-level kernel
    if (k == 1) return ((rows + TV - 1) / TV) * ((cols + TU - 1) / TU);
    unsigned tv, tu;
    if (shape == kTBShape32x32) tb_region<32, 32>(k, rows, cols, &tv, &tu);
    else tb_region<64, 32>(k, rows, cols, &tv, &tu);
    return tv * tu;
}

template <int RV, int RU, bool FRESH>
static hipError_t launch_tb_batch(int k, const BatchTable &t, const dim3 &grid, hipStream_t s)
{
    const dim3 block(kTBThreads);
    switch (k) {
    case 2: hipLaunchKernelGGL((kb_pd_tb<2, RV, RU, FRESH>), grid, block, 0, s, t); break;
    case 3: hipLaunchKernelGGL((kb_pd_tb<3, RV, RU, FRESH>), grid, block, 0, s, t); break;
    case 4: hipLaunchKernelGGL((kb_pd_tb<4, RV, RU, FRESH>), grid, block, 0, s, t); break;
    case 5: hipLaunchKernelGGL((kb_pd_tb<5, RV, RU, FRESH>), grid, block, 0, s, t); break;
    case 6: hipLaunchKernelGGL((kb_pd_tb<6, RV, RU, FRESH>), grid, block, 0, s, t); break;
    case 8: hipLaunchKernelGGL((kb_pd_tb<8, RV, RU, FRESH>), grid, block, 0, s, t); break;
    case 10: hipLaunchKernelGGL((kb_pd_tb<10, RV, RU, FRESH>), grid, block, 0, s, t); break;
    case 12: hipLaunchKernelGGL((kb_pd_tb<12, RV, RU, FRESH>), grid, block, 0, s, t); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int RV, int RU>
static hipError_t launch_tb_batch(int k, bool fresh, const BatchTable &t, const dim3 &grid, hipStream_t s)
{
    return fresh ? launch_tb_batch<RV, RU, true>(k, t, grid, s) : launch_tb_batch<RV, RU, false>(k, t, grid, s);
}

hipError_t launch_pd_batch(int k, int shape, bool fresh, const BatchTable &t, unsigned total_blocks, hipStream_t s)
{
    if (total_blocks == 0 || t.n == 0) return hipSuccess;
    const dim3 grid(total_blocks);
    switch (k) {
    case 0: hipLaunchKernelGGL(kb_pd_level, grid, dim3(BT), 0, s, t); return hipGetLastError();
    case 1: hipLaunchKernelGGL(kb_pd_fused, grid, dim3(TV * TU), 0, s, t); return hipGetLastError();
    default:
        switch (shape) {
        case kTBShape64x32: return launch_tb_batch<64, 32>(k, fresh, t, grid, s);
        case kTBShape32x32: return launch_tb_batch<32, 32>(k, fresh, t, grid, s);
        default: return hipErrorInvalidValue;
        }
    }
}

}  // namespace pdf
