// Fused per-level set-up kernels of the solve path (pdf_solve): the reference's eleven
// set-up launches per level (K2-K8, K13-K15; pdflow_cudalib.cu:439-808, :1018-1167) become four,
// and the two frames' pyramids are built together.  Each output value is computed with the
// same operations as the reference kernel that produces it, so results are bit-identical to
// the stage-by-stage bridges (which remain available through the C-ABI).
//
//   k_level_init      K2 + K3 + K4: zero the level state and, above the coarsest level, filter
//                     the previous level's solution up without materialising the zero-inserted
//                     planes (a zero tap contributes g * 0.f exactly as in the reference).
//   k_rij_gradients   K5 + K6: connectivity weights, then the weighted image gradients, which
//                     need ri_2/rj_2 one pixel back: recomputed here instead of a barrier.
//   k_warp_mu         K7 + K8: warping, then mu_uv (and, on request, the step-size planes).
//   k_filter_motion   K13 + K14 + K15: the weighted median reads the saturated solution of its
//                     3x3 neighbourhood, computed on the fly, and the motion field follows.
//   k_pyramid2        K1 for two frames in one launch per level.
#include <hip/hip_runtime.h>

#include "pdflow_internal.hpp"
#include "pdflow_pixel.hpp"

namespace pdf {
namespace {

constexpr int kBlock = 256;
constexpr int kPyr0PerThread = 4;   // level-0 pixels per thread of the batched pyramid


inline dim3 grid_for(unsigned n) { return dim3((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------------ K2 + K3 + K4
__constant__ int kInitZero[] = {PDF_DU_NEW, PDF_DV_NEW, PDF_DU_ACC, PDF_DV_ACC};
__constant__ int kCarry[10] = {PDF_DU_L, PDF_DV_L, PDF_DW_L, PDF_PD_L, PDF_PUU_L,
                               PDF_PUV_L, PDF_PVU_L, PDF_PVV_L, PDF_PWU_L, PDF_PWV_L};
__constant__ int kUpDst[10] = {PDF_DU_PREV, PDF_DV_PREV, PDF_DW_NEW, PDF_PD, PDF_PUU,
                               PDF_PUV, PDF_PVU, PDF_PVV, PDF_PWU, PDF_PWV};

// 2-D tiles of TR rows x TC columns (one pixel per thread, lanes down a column) for the set-up
// kernels that read a neighbourhood: the neighbourhood is staged through LDS once per tile.  The
// single-pair kernels and the pyramid use 32 x 8 (256 threads); a batch's level-init and median
// filter take any tile of PDF_SETUP_TILES (pdf_batch_set_setup_tile), results unchanged.
constexpr int kTR = 32, kTC = 8;
template <int TR = kTR, int TC = kTC>
__host__ __device__ inline unsigned tiles_of(unsigned rows, unsigned cols)
{
    return ((rows + TR - 1) / TR) * ((cols + TC - 1) / TC);
}

// the coarse patch a tile's 5x5 even taps touch: coarse rows v0/2-1 .. v0/2+TR/2, columns
// u0/2-1 .. u0/2+TC/2 (clamped to the coarse level), at most (TR/2+2) x (TC/2+2), stored with
// row stride CR = TR/2+3; element CR*CC of each plane holds +0 for the taps that fall outside
// the level
template <int TR = kTR, int TC = kTC>
__device__ __forceinline__ void level_init_body(const LevelArgs &a, unsigned tile, int has_coarse,
                                                bool write_const = true)
{
    constexpr int kCR = TR / 2 + 3, kCC = TC / 2 + 3, kZero = kCR * kCC;
    static_assert(kCR * kCC <= TR * TC, "one coarse element per thread and plane");
    static_assert(TR % 2 == 0 && TC % 2 == 0, "tiles start at even rows and columns");
    __shared__ float sC[10][kZero + 1];
    const unsigned rows_i = a.rows, cols_i = a.cols;
    const unsigned tv = (rows_i + TR - 1) / TR;
    const int v0 = (int)(tile % tv) * TR, u0 = (int)(tile / tv) * TC;
    const int v = v0 + (int)threadIdx.x % TR, u = u0 + (int)threadIdx.x / TR;
    const bool on = v < (int)rows_i && u < (int)cols_i;
    const unsigned index = on ? (unsigned)v + (unsigned)u * rows_i : 0u;
    float acc[10];
#pragma unroll
    for (int q = 0; q < 10; q++) acc[q] = 0.f;
    if (has_coarse) {
        const int rc = (int)rows_i / 2, cc = (int)cols_i / 2;
        const int cr0 = v0 >= 2 ? (v0 - 2) / 2 : 0, cr1 = min(rc - 1, (v0 + TR + 1) / 2);
        const int cc0 = u0 >= 2 ? (u0 - 2) / 2 : 0, cc1 = min(cc - 1, (u0 + TC + 1) / 2);
        const int t = (int)threadIdx.x, tr = t % kCR, tc = t / kCR;
        if (t < kCR * kCC && tr <= cr1 - cr0 && tc <= cc1 - cc0) {
            const unsigned src = (unsigned)(cr0 + tr) + (unsigned)(cc0 + tc) * (unsigned)rc;
#pragma unroll
            for (int q = 0; q < 10; q++) sC[q][t] = gptr(a.f[kCarry[q]])[src];
        }
        if (t < 10) sC[t][kZero] = 0.f;
        __syncthreads();
        if (on) {
            // .cu:460-569: the zero-inserted plane holds the coarse value at even (v, u) (du, dv
            // doubled) and +0 elsewhere.  A running sum that starts at +0 is never -0 under
            // round-to-nearest, and adding +0 or -0 to it leaves it unchanged: so only the even
            // taps are visited, in the reference's (k outer, l inner) order, and a tap outside
            // the level (or the third tap of an odd row or column) reads the +0 element, whose
            // product with the finite weight is such a zero.  Rows and columns of one parity
            // visit k (l) = -2, 0, 2, the other -1, 1: the loops are unrolled over the three
            // tap slots and each lane picks its weight from the four parity classes, read once
            // per workgroup, so every lane runs the same instructions.
            const bool inner = (v > 1) && (v < (int)rows_i - 2) && (u > 1) && (u < (int)cols_i - 2);
            const int vo = v & 1, uo = u & 1;
            // the 25 weights as wave-uniform values, so that a per-lane choice between two of
            // them selects registers rather than addresses
            float gw[25];
#pragma unroll
            for (int n = 0; n < 25; n++)
                gw[n] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(a.g[n])));
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int k = 2 * j - 2 + vo, vv = v + k;
                const bool vin = (j < 2 || vo == 0) && vv >= 0 && vv < (int)rows_i;
                const int ro = min(max((vv >> 1) - cr0, 0), kCR - 1);
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const int l = 2 * i - 2 + uo, uu = u + l;
                    const bool tin = vin && (i < 2 || uo == 0) && uu >= 0 && uu < (int)cols_i;
                    const int co = min(max((uu >> 1) - cc0, 0), kCC - 1);
                    const int ci = tin ? ro + co * kCR : kZero;
                    // weights of tap slot (j, i) for even/odd v and u (k, l clamped to 2 where
                    // the slot does not exist; its product is zero there)
                    const int ke = 2 * j - 2, ko = min(2 * j - 1, 2);
                    const int le = 2 * i - 2, lo = min(2 * i - 1, 2);
                    const float gv_e = uo ? gw[12 + ke + 5 * lo] : gw[12 + ke + 5 * le];
                    const float gv_o = uo ? gw[12 + ko + 5 * lo] : gw[12 + ko + 5 * le];
                    const float g = 4.f * (vo ? gv_o : gv_e);
#pragma unroll
                    for (int q = 0; q < 10; q++) {
                        const float cval = sC[q][ci];
                        acc[q] += g * (q < 2 ? 2.f * cval : cval);
                    }
                }
            }
            if (!inner) {
                // border weight: 1 - sum of g over the out-of-range taps (.cu:515-542)
                float aw = 1.f;
#pragma unroll
                for (int k = -2; k < 3; k++)
#pragma unroll
                    for (int l = -2; l < 3; l++) {
                        const int vv = v + k, uu = u + l;
                        if (vv < 0 || vv >= (int)rows_i || uu < 0 || uu >= (int)cols_i)
                            aw -= 0.25f * (4.f * a.g[12 + k + 5 * l]);
                    }
                const float inv = 1.f / aw;
#pragma unroll
                for (int q = 0; q < 10; q++) acc[q] *= inv;
            }
        }
    }
    if (!on) return;
#pragma unroll
    for (int q = 0; q < 10; q++) gptr(a.f[kUpDst[q]])[index] = acc[q];
    if (write_const) {
        // .cu:447-454: du_new, dv_new, du_acc, dv_acc start at 0 on every level
#pragma unroll
        for (int k = 0; k < 4; k++) gptr(a.f[kInitZero[k]])[index] = 0.f;
        gptr(a.f[PDF_DW_ACC])[index] = acc[2];
    }
}

// ------------------------------------------------------------------ warm start
// Replaces K2-K4 at the level a solve starts from: du_prev, dv_prev (pixels at this level)
// and dw (metres) come from three given planes, every dual and increment starts at 0
// (the injection point of the learned initialisation; no reference counterpart).
__device__ __forceinline__ void warm_init_body(const LevelArgs &a, unsigned block)
{
    const unsigned index = block * kBlock + threadIdx.x;
    const unsigned P = a.rows * a.cols;
    if (index >= P) return;
    const float du = gptr(a.warm)[index], dv = gptr(a.warm)[P + index], dw = gptr(a.warm)[2 * P + index];
#pragma unroll
    for (int k = 0; k < 4; k++) gptr(a.f[kInitZero[k]])[index] = 0.f;
    gptr(a.f[PDF_DU_PREV])[index] = du;
    gptr(a.f[PDF_DV_PREV])[index] = dv;
    gptr(a.f[PDF_DW_NEW])[index] = dw;
    gptr(a.f[PDF_DW_ACC])[index] = dw;
#pragma unroll
    for (int q = 3; q < 10; q++) gptr(a.f[kUpDst[q]])[index] = 0.f;   // pd and the six TV duals
}

// ------------------------------------------------------------------ K5 + K6
__device__ __forceinline__ float rweight(float d1, float d2)
{
    return (fabsf(d1) + fabsf(d2) > 0.f) ? 2.f * rhyp(d1, d2) : 1.f;     // .cu:760-778
}

__device__ __forceinline__ void rij_gradients_body(const LevelArgs &a, unsigned block)
{
    const unsigned index = block * kBlock + threadIdx.x;
    const unsigned rows_i = a.rows, cols_i = a.cols, r = rows_i;
    if (index >= rows_i * cols_i) return;
    const unsigned v = index % rows_i, u = index / rows_i;
    const bool uf = u == 0, ul = u == cols_i - 1, vf = v == 0, vl = v == rows_i - 1;
    // Every value is loaded up front, so the loads are in flight together.  A neighbour index
    // that would leave the plane is clamped to the pixel itself: the load stays in bounds, and
    // the border cases below never use that value.
    const unsigned iu1 = ul ? index : index + r, iu0 = uf ? index : index - r;
    const unsigned iv1 = vl ? index : index + 1, iv0 = vf ? index : index - 1;
    const auto X = gptr(a.X), Y = gptr(a.Y), Z = gptr(a.Z), I = gptr(a.I);
    const auto Xo = gptr(a.Xold), Yo = gptr(a.Yold), Zo = gptr(a.Zold);
    const float x_c = X[index], x_u1 = X[iu1], x_u0 = X[iu0];
    const float y_c = Y[index], y_v1 = Y[iv1], y_v0 = Y[iv0];
    const float z_c = Z[index], z_u1 = Z[iu1], z_u0 = Z[iu0], z_v1 = Z[iv1], z_v0 = Z[iv0];
    const float i_c = I[index], i_u1 = I[iu1], i_u0 = I[iu0], i_v1 = I[iv1], i_v0 = I[iv0];
    const float xo_c = Xo[index], xo_u1 = Xo[iu1], yo_c = Yo[index], yo_v1 = Yo[iv1];
    const float zo_c = Zo[index], zo_u1 = Zo[iu1], zo_v1 = Zo[iv1];
    // K5 (.cu:725-779): the differences towards u+1 and v+1 are 0 in the last column / row
    const float dxu = ul ? 0.f : xo_u1 - xo_c, dzu = ul ? 0.f : zo_u1 - zo_c;
    const float dxu2 = ul ? 0.f : x_u1 - x_c, dzu2 = ul ? 0.f : z_u1 - z_c;
    const float dyv = vl ? 0.f : yo_v1 - yo_c, dzv = vl ? 0.f : zo_v1 - zo_c;
    const float dyv2 = vl ? 0.f : y_v1 - y_c, dzv2 = vl ? 0.f : z_v1 - z_c;
    const float ri2 = rweight(dxu2, dzu2), rj2 = rweight(dyv2, dzv2);
    const float ri = rweight(dxu, dzu), rj = rweight(dyv, dzv);
    // K6 (.cu:574-627); ri_2 at (v, u-1) and rj_2 at (v-1, u) recomputed (u-1 and v-1 are
    // never the last column/row there, so their differences always exist)
    float dcu, ddu, dcv, ddv;
    if (uf) {
        dcu = i_u1 - i_c;
        ddu = z_u1 - z_c;
    } else if (ul) {
        dcu = i_c - i_u0;
        ddu = z_c - z_u0;
    } else {
        const float rp = rweight(x_c - x_u0, z_c - z_u0);
        dcu = wdiff(ri2, rp, i_u1, i_c, i_u0);
        ddu = (z_c > 0.f) ? wdiff(ri2, rp, z_u1, z_c, z_u0) : 0.f;
    }
    if (vf) {
        dcv = i_v1 - i_c;
        ddv = z_v1 - z_c;
    } else if (vl) {
        dcv = i_c - i_v0;
        ddv = z_c - z_v0;
    } else {
        const float rp = rweight(y_c - y_v0, z_c - z_v0);
        dcv = wdiff(rj2, rp, i_v1, i_c, i_v0);
        ddv = (z_c > 0.f) ? wdiff(rj2, rp, z_v1, z_c, z_v0) : 0.f;
    }
    gptr(a.f[PDF_RI])[index] = ri;
    gptr(a.f[PDF_RJ])[index] = rj;
    gptr(a.f[PDF_RI_2])[index] = ri2;
    gptr(a.f[PDF_RJ_2])[index] = rj2;
    gptr(a.f[PDF_DCU_AUX])[index] = dcu;
    gptr(a.f[PDF_DDU_AUX])[index] = ddu;
    gptr(a.f[PDF_DCV_AUX])[index] = dcv;
    gptr(a.f[PDF_DDV_AUX])[index] = ddv;
}

// ------------------------------------------------------------------ K7 + K8
__device__ __forceinline__ void warp_mu_body(const LevelArgs &a, unsigned block, int write_steps)
{
    const unsigned index = block * kBlock + threadIdx.x;
    const unsigned rows_i = a.rows, cols_i = a.cols;
    if (index >= rows_i * cols_i) return;
    const unsigned v = index % rows_i, u = index / rows_i;
    // K7 (.cu:629-652)
    const float iu = (float)u + gptr(a.f[PDF_DU_PREV])[index];
    const float iv = (float)v + gptr(a.f[PDF_DV_PREV])[index];
    const Tap t = make_tap(rows_i, cols_i, iu, iv);
    // the 24 corners of the six gathered planes and the two frame-1 values are all issued
    // before any of them is used; the output planes and mu are read before that point too
    const Corners cI = corners(a.I, rows_i, t), cCU = corners(a.f[PDF_DCU_AUX], rows_i, t),
                  cCV = corners(a.f[PDF_DCV_AUX], rows_i, t), cZ = corners(a.Z, rows_i, t),
                  cDU = corners(a.f[PDF_DDU_AUX], rows_i, t), cDV = corners(a.f[PDF_DDV_AUX], rows_i, t);
    const float iold = gptr(a.Iold)[index], zold = gptr(a.Zold)[index];
    float *const o_ct = a.f[PDF_DCT], *const o_cu = a.f[PDF_DCU], *const o_cv = a.f[PDF_DCV];
    float *const o_dt = a.f[PDF_DDT], *const o_du = a.f[PDF_DDU], *const o_dv = a.f[PDF_DDV];
    float *const o_mu = a.f[PDF_MU_UV];
    const float mu = a.mu;
    issued(cI); issued(cCU); issued(cCV); issued(cZ); issued(cDU); issued(cDV);
    const float dct = bilinear(cI, t) - iold;
    const float dcu = bilinear(cCU, t);
    const float dcv = bilinear(cCV, t);
    const float wd = bilinear_depth(cZ, t);
    const float ddt = (wd > 0.f) ? wd - zold : 0.f;
    const float ddu = bilinear(cDU, t);
    const float ddv = bilinear(cDV, t);
    gptr(o_ct)[index] = dct;
    gptr(o_cu)[index] = dcu;
    gptr(o_cv)[index] = dcv;
    gptr(o_dt)[index] = ddt;
    gptr(o_du)[index] = ddu;
    gptr(o_dv)[index] = ddv;
    // K8 computeMu (.cu:781-784)
    const float mu_uv = mu / (1.f + 1000.f * (ddu * ddu + ddv * ddv + ddt * ddt));
    gptr(o_mu)[index] = mu_uv;
    if (write_steps) {
        // K8 computeStepSizes (.cu:786-808), needed only by the unfused iteration bridges
        const float ri = gptr(a.f[PDF_RI])[index], rj = gptr(a.f[PDF_RJ])[index];
        gptr(a.f[PDF_SIGMA_PD])[index] = sigma_pd_of(mu_uv, ddu, ddv);
        gptr(a.f[PDF_SIGMA_PUVX])[index] = sigma_p_of(a.lambda_i, ri);
        gptr(a.f[PDF_SIGMA_PUVY])[index] = sigma_p_of(a.lambda_i, rj);
        gptr(a.f[PDF_SIGMA_PWX])[index] = 0.5f / (ri * a.lambda_d + 1e-10f);
        gptr(a.f[PDF_SIGMA_PWY])[index] = 0.5f / (rj * a.lambda_d + 1e-10f);
        float acu_r = ri + rj;
        if (u > 0) acu_r += gptr(a.f[PDF_RI])[index - rows_i];
        if (v > 0) acu_r += gptr(a.f[PDF_RJ])[index - 1];
        gptr(a.f[PDF_TAU_U])[index] = tau_uv_of(mu_uv, ddu, a.lambda_i, acu_r);
        gptr(a.f[PDF_TAU_V])[index] = tau_uv_of(mu_uv, ddv, a.lambda_i, acu_r);
        gptr(a.f[PDF_TAU_W])[index] = tau_w_of(mu_uv, a.lambda_d, acu_r);
    }
}

// ------------------------------------------------------------------ K13 + K14 + K15
struct Sat {
    float du, dv, dw;
};

__device__ __forceinline__ Sat saturated(const LevelArgs &a, unsigned i)
{
    // saturateVariables (.cu:1018-1038) of pixel i
    float du = gptr(a.f[PDF_DU_NEW])[i], dv = gptr(a.f[PDF_DV_NEW])[i], dw = gptr(a.f[PDF_DW_NEW])[i];
    if (du > 1.f) du = 1.f; else if (du < -1.f) du = -1.f;
    if (dv > 1.f) dv = 1.f; else if (dv < -1.f) dv = -1.f;
    if (gptr(a.Zold)[i] == 0.f) dw = 0.f;
    return {du + gptr(a.f[PDF_DU_PREV])[i], dv + gptr(a.f[PDF_DV_PREV])[i], dw};
}

// carry_duals: copy pd and the six duals into the *_l carry planes that the next level's
// upsampling reads (.cu:1134-1147); a pair's finest level has no next level
template <int TR = kTR, int TC = kTC>
__device__ __forceinline__ void filter_motion_body(const LevelArgs &a, unsigned tile, bool carry_duals = true)
{
    // the tile plus a one-pixel halo: saturated solution, frame-1 depth and ddt of each pixel,
    // computed once per pixel and read by the (up to) 9 pixels whose window contains it
    constexpr int HR = TR + 2, HC = TC + 2;
    __shared__ float sU[HR * HC], sV[HR * HC], sW[HR * HC], sZ[HR * HC], sT[HR * HC];
    const unsigned rows_i = a.rows, cols_i = a.cols;
    const unsigned tv = (rows_i + TR - 1) / TR;
    const int v0 = (int)(tile % tv) * TR, u0 = (int)(tile / tv) * TC;
    for (int e = (int)threadIdx.x; e < HR * HC; e += TR * TC) {
        const int vv = v0 - 1 + e % HR, uu = u0 - 1 + e / HR;
        if (vv >= 0 && vv < (int)rows_i && uu >= 0 && uu < (int)cols_i) {
            const unsigned il = (unsigned)vv + (unsigned)uu * rows_i;
            const Sat f = saturated(a, il);
            sU[e] = f.du;
            sV[e] = f.dv;
            sW[e] = f.dw;
            sZ[e] = gptr(a.Zold)[il];
            sT[e] = gptr(a.f[PDF_DDT])[il];
        }
    }
    __syncthreads();
    const int lv = (int)threadIdx.x % TR, lu = (int)threadIdx.x / TR;
    const int v = v0 + lv, u = u0 + lu;
    if (v >= (int)rows_i || u >= (int)cols_i) return;
    const unsigned index = (unsigned)v + (unsigned)u * rows_i;
    const int hc = (lv + 1) + (lu + 1) * HR;            // this pixel in the halo tile
    const float d0 = sZ[hc];
    float ou, ov, ow;
    if (d0 > 0.f) {
        // .cu:1057-1132: taps in the reference's order (k = v offset outer, l = u offset inner)
        // the weight 1 / (1 + 5 dz^2 + 10 dt^2) has a denominator >= 1 (or NaN), which the short
        // exact reciprocal covers up to 2^124 and the full division beyond
        auto weight = [&](int h) {
            const float dz = d0 - sZ[h], dt = sT[h];
            return div_half_or_one(1.f, 1.f + 5.f * (dz * dz) + 10.f * (dt * dt));
        };
        float ku[9], kv[9], kw[9], wu[9], wv[9], ww[9];
        const bool interior = v > 0 && v < (int)rows_i - 1 && u > 0 && u < (int)cols_i - 1;
        if (interior) {
            // all nine taps exist: tap s goes to slot s
#pragma unroll
            for (int k = -1; k < 2; k++)
#pragma unroll
                for (int l = -1; l < 2; l++) {
                    const int s = (k + 1) * 3 + (l + 1), h = hc + k + l * HR;
                    ku[s] = sU[h]; kv[s] = sV[h]; kw[s] = sW[h];
                    wu[s] = wv[s] = ww[s] = weight(h);
                }
        } else {
            // the existing taps fill the first slots in order; the others carry key +inf and
            // weight 0
            int n = 0;
#pragma unroll
            for (int s = 0; s < 9; s++) {
                ku[s] = kv[s] = kw[s] = __builtin_inff();
                wu[s] = wv[s] = ww[s] = 0.f;
            }
#pragma unroll
            for (int k = -1; k < 2; k++)
#pragma unroll
                for (int l = -1; l < 2; l++) {
                    const int r = v + k, c = u + l;
                    const bool ok = !(r < 0 || r >= (int)rows_i || c < 0 || c >= (int)cols_i);
                    const int h = ok ? hc + k + l * HR : hc;
                    const float pres = weight(h);
                    const float fu = sU[h], fv = sV[h], fw = sW[h];
#pragma unroll
                    for (int s = 0; s < 9; s++)
                        if (ok && s == n) {
                            ku[s] = fu; kv[s] = fv; kw[s] = fw;
                            wu[s] = pres; wv[s] = pres; ww[s] = pres;
                        }
                    n += ok ? 1 : 0;
                }
        }
        sort9x3(ku, wu, kv, wv, kw, ww);
        float cu[9], cv[9], cw[9];
        cu[0] = wu[0]; cv[0] = wv[0]; cw[0] = ww[0];
#pragma unroll
        for (int i = 1; i < 9; i++) {
            cu[i] = cu[i - 1] + wu[i];
            cv[i] = cv[i - 1] + wv[i];
            cw[i] = cw[i - 1] + ww[i];
        }
        const float med = 0.5f * cu[8];
        ou = wmedian_pick(ku, cu, med);
        ov = wmedian_pick(kv, cv, med);
        ow = wmedian_pick(kw, cw, med);
    } else {
        ou = sU[hc]; ov = sV[hc]; ow = sW[hc];
    }
    gptr(a.f[PDF_DU_L])[index] = ou;
    gptr(a.f[PDF_DV_L])[index] = ov;
    gptr(a.f[PDF_DW_L])[index] = ow;
    if (carry_duals) {
        gptr(a.f[PDF_PD_L])[index] = gptr(a.f[PDF_PD])[index];
        gptr(a.f[PDF_PUU_L])[index] = gptr(a.f[PDF_PUU])[index];
        gptr(a.f[PDF_PUV_L])[index] = gptr(a.f[PDF_PUV])[index];
        gptr(a.f[PDF_PVU_L])[index] = gptr(a.f[PDF_PVU])[index];
        gptr(a.f[PDF_PVV_L])[index] = gptr(a.f[PDF_PVV])[index];
        gptr(a.f[PDF_PWU_L])[index] = gptr(a.f[PDF_PWU])[index];
        gptr(a.f[PDF_PWV_L])[index] = gptr(a.f[PDF_PWV])[index];
    }
    // K15 (.cu:1150-1167)
    if (d0 > 0.f) {
        gptr(a.f[PDF_DX])[index] = ow;
        gptr(a.f[PDF_DY])[index] = d0 * ou * a.inv_f + ow * gptr(a.Xold)[index] / d0;
        gptr(a.f[PDF_DZ])[index] = d0 * ov * a.inv_f + ow * gptr(a.Yold)[index] / d0;
    } else {
        gptr(a.f[PDF_DX])[index] = 0.f;
        gptr(a.f[PDF_DY])[index] = 0.f;
        gptr(a.f[PDF_DZ])[index] = 0.f;
    }
}

// ------------------------------------------------------------------ K1 for two frames
// One 32x8 tile of level rows_i x cols_i per block.  Above level 0 the tile's 5x5 neighbourhoods
// in the finer level (rows 2*v0-2 .. 2*v0+64, columns 2*u0-2 .. 2*u0+16) are staged through LDS
// first, by independent loads that run down the columns; the taps then read LDS.
constexpr int kPR = 2 * kTR + 3, kPC = 2 * kTC + 3;   // 67 x 19 patch
constexpr int kPatch = kPR * kPC;

// Level 0 is a copy with PX0 pixels per thread: block `tile` covers pixels
// [tile * PX0 * kBlock, (tile + 1) * PX0 * kBlock) in column-major order, one contiguous run of
// kBlock pixels per pass, so that a wave's accesses are 256 contiguous bytes.
template <int PX0>
__device__ __forceinline__ void pyramid_body(const PyrArgs &p, unsigned tile)
{
    __shared__ float sCol[kPatch], sDep[kPatch];
    const unsigned rows_i = p.rows_i, cols_i = p.cols_i;
    const float max_depth_dif = 0.1f;
    const float disp_u = 0.5f * (float)(cols_i - 1);
    const float disp_v = 0.5f * (float)(rows_i - 1);
    if (p.level == 0) {
        // A batch's level-0 colour and depth planes are its input planes themselves, and only
        // xx, yy are written.  The loads of all passes are issued first (at an index clamped
        // into the plane; a pass past the end stores nothing).
        const unsigned P = rows_i * cols_i;
        float depth[PX0], colour[PX0];
        const bool copy_col = p.col != p.colour_wf, copy_dep = p.dep != p.depth_wf;
#pragma unroll
        for (int i = 0; i < PX0; i++) {
            const unsigned index = min((tile * PX0 + i) * kBlock + threadIdx.x, P - 1);
            depth[i] = gptr(p.depth_wf)[index];
            colour[i] = copy_col ? gptr(p.colour_wf)[index] : 0.f;
        }
#pragma unroll
        for (int i = 0; i < PX0; i++) {
            const unsigned index = (tile * PX0 + i) * kBlock + threadIdx.x;
            if (index >= P) break;
            const unsigned v = index % rows_i, u = index / rows_i;
            if (copy_col) gptr(p.col)[index] = colour[i];
            if (copy_dep) gptr(p.dep)[index] = depth[i];
            gptr(p.xx)[index] = ((float)u - disp_u) * depth[i] * p.inv_f_i;
            gptr(p.yy)[index] = ((float)v - disp_v) * depth[i] * p.inv_f_i;
        }
        return;
    }
    const unsigned tv = (rows_i + kTR - 1) / kTR;
    const int v0 = (int)(tile % tv) * kTR, u0 = (int)(tile / tv) * kTC;
    const int v = v0 + (int)threadIdx.x % kTR, u = u0 + (int)threadIdx.x / kTR;
    const bool on = v < (int)rows_i && u < (int)cols_i;
    const unsigned index = (unsigned)v + (unsigned)u * rows_i;   // read only when on
    const int pr = 2 * (int)rows_i, pc = 2 * (int)cols_i;   // the finer level
    const int r0 = 2 * v0 - 2, c0 = 2 * u0 - 2;
    // every thread's patch loads are issued before the first is stored: coordinates outside
    // the finer level are clamped into it, and the cells they fill are never read (an inner
    // pixel's taps are all in range, a border pixel skips the taps outside)
    constexpr int kLoads = (kPatch + kBlock - 1) / kBlock;
    float lc[kLoads], ld[kLoads];
#pragma unroll
    for (int i = 0; i < kLoads; i++) {
        const int e = (int)threadIdx.x + i * kBlock;
        const int vv = min(max(r0 + e % kPR, 0), pr - 1), uu = min(max(c0 + e / kPR, 0), pc - 1);
        const unsigned ip = (unsigned)vv + (unsigned)uu * (unsigned)pr;
        lc[i] = gptr(p.colp)[ip];
        ld[i] = gptr(p.depp)[ip];
    }
#pragma unroll
    for (int i = 0; i < kLoads; i++) {
        const int e = (int)threadIdx.x + i * kBlock;
        if (e < kPatch) {
            sCol[e] = lc[i];
            sDep[e] = ld[i];
        }
    }
    __syncthreads();
    if (!on) return;
    // .cu:357-424, taps in the reference's (k outer, l inner) order
    const int c = (2 * (v - v0) + 2) + (2 * (u - u0) + 2) * kPR;   // (2v, 2u) in the patch
    const float dcen = sDep[c];
    float sumd = 0.f, sumc = 0.f, wd = 0.f, wc = 0.f;
    const bool inner = (v > 0) && (v < (int)rows_i - 1) && (u > 0) && (u < (int)cols_i - 1);
    if (inner) {
        // every tap is in range.  A rejected depth tap adds +0 to wd and sumd instead of
        // being skipped.  Both sums start at +0, and under round-to-nearest a sum only reaches
        // zero by exact cancellation, which gives +0: they are never -0, so adding +0 leaves
        // them unchanged whatever the mask and depths.
#pragma unroll
        for (int k = -2; k < 3; k++)
#pragma unroll
            for (int l = -2; l < 3; l++) {
                const float g = p.g[12 + k + 5 * l];
                const int s = c + k + l * kPR;
                sumc += g * sCol[s];
                const float d = sDep[s];
                const bool take = (d > 0.f) && (fabsf(d - dcen) < max_depth_dif);
                const float w = g * (max_depth_dif - fabsf(d - dcen));
                wd += take ? w : 0.f;
                sumd += take ? w * d : 0.f;
            }
        gptr(p.col)[index] = sumc;
    } else {
        for (int k = -2; k < 3; k++)
            for (int l = -2; l < 3; l++) {
                const int vv = 2 * v + k, uu = 2 * u + l;
                if (!(vv >= 0 && vv < pr && uu >= 0 && uu < pc)) continue;
                const float g = p.g[12 + k + 5 * l];
                const int s = c + k + l * kPR;
                sumc += g * sCol[s];
                wc += g;
                const float d = sDep[s];
                if ((d > 0.f) && (fabsf(d - dcen) < max_depth_dif)) {
                    const float w = g * (max_depth_dif - fabsf(d - dcen));
                    wd += w;
                    sumd += w * d;
                }
            }
        gptr(p.col)[index] = sumc / wc;
    }
    const float depth = (sumd > 0.f) ? sumd / wd : 0.f;
    gptr(p.dep)[index] = depth;
    gptr(p.xx)[index] = ((float)u - disp_u) * depth * p.inv_f_i;
    gptr(p.yy)[index] = ((float)v - disp_v) * depth * p.inv_f_i;
}

// ------------------------------------------------------------------ entry points
// The batched set-up kernels take their blocks in XCD-contiguous order (xcd_swizzle), so that
// blocks that share halo rows, neighbouring columns or gathered lines run on one XCD and meet
// in its L2; the blocks of a pair stay consecutive in that order.
__device__ __forceinline__ BatchItem setup_item(const BatchTable &t)
{
    return batch_item(t, xcd_swizzle(blockIdx.x, gridDim.x));
}

// Single level: arguments by value.  Batched: one launch over many pairs, each block finds
// its pair's arguments in a device table (see BatchTable).
__global__ void __launch_bounds__(kBlock) k_level_init(LevelArgs a, int has_coarse) { level_init_body(a, blockIdx.x, has_coarse); }
__global__ void __launch_bounds__(kBlock) k_rij_gradients(LevelArgs a) { rij_gradients_body(a, blockIdx.x); }
// a level's level-init tiles and rij/gradients blocks in one launch: they read and write disjoint
// planes, and a single pair's small levels pay about a launch's latency for each kernel
__global__ void __launch_bounds__(kBlock) k_level_init_rij(LevelArgs a, int has_coarse, unsigned init_tiles)
{
    if (blockIdx.x < init_tiles)
        level_init_body(a, blockIdx.x, has_coarse);
    else
        rij_gradients_body(a, blockIdx.x - init_tiles);
}
__global__ void __launch_bounds__(kBlock) k_warm_init(LevelArgs a) { warm_init_body(a, blockIdx.x); }
__global__ void __launch_bounds__(kBlock) k_warp_mu(LevelArgs a, int write_steps) { warp_mu_body(a, blockIdx.x, write_steps); }
__global__ void __launch_bounds__(kBlock) k_filter_motion(LevelArgs a) { filter_motion_body(a, blockIdx.x); }
__global__ void __launch_bounds__(kBlock) k_pyramid2(Pyr2Args a) { pyramid_body<1>(a.frame[blockIdx.y], blockIdx.x); }

template <int TR, int TC>
__global__ void __launch_bounds__(TR * TC) kb_level_init(BatchTable t, int flag)
{
    const BatchItem it = setup_item(t);
    level_init_body<TR, TC>(static_cast<const LevelArgs *>(t.args)[it.item], it.block, flag & kInitCoarse,
                            (flag & kInitSkipConst) == 0);
}
__global__ void __launch_bounds__(kBlock) kb_warm_init(BatchTable t)
{
    const BatchItem it = setup_item(t);
    warm_init_body(static_cast<const LevelArgs *>(t.args)[it.item], it.block);
}
__global__ void __launch_bounds__(kBlock) kb_rij_gradients(BatchTable t)
{
    const BatchItem it = setup_item(t);
    rij_gradients_body(static_cast<const LevelArgs *>(t.args)[it.item], it.block);
}
__global__ void __launch_bounds__(kBlock) kb_warp_mu(BatchTable t)
{
    const BatchItem it = setup_item(t);
    warp_mu_body(static_cast<const LevelArgs *>(t.args)[it.item], it.block, 0);
}
// flag kFilterNoCarry: the duals are not copied into the *_l planes
template <int TR, int TC>
__global__ void __launch_bounds__(TR * TC) kb_filter_motion(BatchTable t, int flag)
{
    const BatchItem it = setup_item(t);
    filter_motion_body<TR, TC>(static_cast<const LevelArgs *>(t.args)[it.item], it.block,
                               (flag & kFilterNoCarry) == 0);
}
__global__ void __launch_bounds__(kBlock) kb_pyramid(BatchTable t)
{
    const BatchItem it = setup_item(t);
    pyramid_body<kPyr0PerThread>(static_cast<const PyrArgs *>(t.args)[it.item], it.block);
}

}  // namespace

hipError_t launch_level_init(const LevelArgs &a, bool has_coarse, hipStream_t s)
{
    if (a.rows == 0 || a.cols == 0) return hipSuccess;
    hipLaunchKernelGGL(k_level_init, dim3(tiles_of(a.rows, a.cols)), dim3(kBlock), 0, s, a, has_coarse ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_warm_init(const LevelArgs &a, hipStream_t s)
{
    const unsigned n = a.rows * a.cols;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_warm_init, grid_for(n), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_level_init_rij(const LevelArgs &a, bool has_coarse, hipStream_t s)
{
    const unsigned n = a.rows * a.cols;
    if (n == 0) return hipSuccess;
    const unsigned tiles = tiles_of(a.rows, a.cols);
    hipLaunchKernelGGL(k_level_init_rij, dim3(tiles + grid_for(n).x), dim3(kBlock), 0, s, a, has_coarse ? 1 : 0,
                       tiles);
    return hipGetLastError();
}

hipError_t launch_rij_gradients(const LevelArgs &a, hipStream_t s)
{
    const unsigned n = a.rows * a.cols;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rij_gradients, grid_for(n), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_warp_mu(const LevelArgs &a, bool write_steps, hipStream_t s)
{
    const unsigned n = a.rows * a.cols;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_warp_mu, grid_for(n), dim3(kBlock), 0, s, a, write_steps ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_filter_motion(const LevelArgs &a, hipStream_t s)
{
    if (a.rows == 0 || a.cols == 0) return hipSuccess;
    hipLaunchKernelGGL(k_filter_motion, dim3(tiles_of(a.rows, a.cols)), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pyramid2(const Pyr2Args &a, hipStream_t s)
{
    const unsigned rows = a.frame[0].rows_i, cols = a.frame[0].cols_i;
    if (rows == 0 || cols == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pyramid2, dim3(tiles_of(rows, cols), 2), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

}  // namespace pdf

namespace pdf {

bool setup_tile_supported(int tile) { return tile >= 0 && tile < kNumSetupTiles; }

hipError_t launch_batch(BatchKernel k, const BatchTable &t, unsigned total_blocks, int flag, int tile,
                        hipStream_t s)
{
    if (total_blocks == 0 || t.n == 0) return hipSuccess;
    const dim3 grid(total_blocks), block(kBlock);
    switch (k) {
    case BatchKernel::LevelInit:
        switch (tile) {
#define PDF_LAUNCH_TILE(id, R, C) \
        case id: hipLaunchKernelGGL((kb_level_init<R, C>), grid, dim3(R * C), 0, s, t, flag); break;
        PDF_SETUP_TILES(PDF_LAUNCH_TILE)
#undef PDF_LAUNCH_TILE
        default: return hipErrorInvalidValue;
        }
        break;
    case BatchKernel::RijGradients: hipLaunchKernelGGL(kb_rij_gradients, grid, block, 0, s, t); break;
    case BatchKernel::WarpMu: hipLaunchKernelGGL(kb_warp_mu, grid, block, 0, s, t); break;
    case BatchKernel::FilterMotion:
        switch (tile) {
#define PDF_LAUNCH_TILE(id, R, C) \
        case id: hipLaunchKernelGGL((kb_filter_motion<R, C>), grid, dim3(R * C), 0, s, t, flag); break;
        PDF_SETUP_TILES(PDF_LAUNCH_TILE)
#undef PDF_LAUNCH_TILE
        default: return hipErrorInvalidValue;
        }
        break;
    case BatchKernel::Pyramid: hipLaunchKernelGGL(kb_pyramid, grid, block, 0, s, t); break;
    case BatchKernel::WarmInit: hipLaunchKernelGGL(kb_warm_init, grid, block, 0, s, t); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

unsigned batch_blocks_per_level(unsigned pixels) { return (pixels + kBlock - 1) / kBlock; }

unsigned batch_pyramid_blocks(unsigned level, unsigned rows, unsigned cols)
{
    if (level == 0) return (rows * cols + kPyr0PerThread * kBlock - 1) / (kPyr0PerThread * kBlock);
    return tiles_of(rows, cols);
}

unsigned batch_blocks(BatchKernel k, unsigned rows, unsigned cols, int tile)
{
    if (k == BatchKernel::LevelInit || k == BatchKernel::FilterMotion) {
        switch (tile) {
#define PDF_TILE_COUNT(id, R, C) \
        case id: return tiles_of<R, C>(rows, cols);
        PDF_SETUP_TILES(PDF_TILE_COUNT)
#undef PDF_TILE_COUNT
        default: return 0;
        }
    }
    if (k == BatchKernel::Pyramid) return tiles_of(rows, cols);
    return batch_blocks_per_level(rows * cols);
}

}  // namespace pdf
