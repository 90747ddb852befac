/* CPU oracle for the PD-Flow plain solve -- TEST INFRASTRUCTURE, the checker only.
 * Interface: pdflow_oracle.h.  Citations are to the reference tree
 * (pdflow_cudalib.cu = ".cu", scene_flow_impair.cpp = "impair.cpp").
 *
 * Arithmetic conventions (each also used by the HIP library, so both produce the same
 * floats):
 *   - FP32 everywhere, evaluated left to right as written in the reference, built with
 *     -ffp-contract=off (no fused multiply-add);
 *   - rhypotf(a,b) is 1/sqrtf(a*a+b*b), the form the reference states next to each call
 *     (.cu:761, :849); fdividef(a,b) is a/b (IEEE division);
 *   - __float2int_ru/_rd/_rn are ceilf/floorf/rintf (round half to even), lrintf and
 *     lroundf are only applied to integral values there and are exact;
 *   - powf(x, 2.f) in the median weights (.cu:1075) is x*x;
 *   - abs() on floats (.cu:791, :805-806) is the float overload, fabsf;
 *   - tan(0.5f*fovh) (.cu:428, :1152) is supplied by the caller as one float.
 * These are the places where CUDA's own intrinsics could differ by an ulp; the HIP
 * library follows the same conventions, so GPU-vs-oracle comparisons are exact
 * up to compiler reassociation, which neither side allows.
 *
 * Pinned by tests/test_oracle_*.py: closed forms (pyramid of constants, x/y
 * back-projection, Rij of a fronto-parallel plane, gradient of a ramp, warping of
 * integer shifts, step-size identities, motion-field formula), the adjoint identity
 * between gradient and divergence, the dual projection invariants, the 3-case prox,
 * zero motion giving exactly zero flow, brute-force bilinear interpolation and
 * brute-force weighted medians on small inputs, and the schedule table.
 */
#include "pdflow_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct orc_ctx {
    unsigned rows, cols, ctf_levels, cam_mode, width, height, pyr_levels;
    float lambda_i, lambda_d, mu, fovh, fovv, tan_half;
    float g_mask[25];
    unsigned rows_i, cols_i, level, level_image;
    size_t cap;                                   /* capacity of every plane */
    float *colour_wf, *depth_wf;
    float *pyr[PDF_NUM_PYR_KINDS][PDF_MAX_PYR_LEVELS];
    float *f[PDF_NUM_FIELDS];
    int nthreads;
};

#define F(id) (c->f[id])

static float rhyp(float a, float b) { return 1.f / sqrtf(a * a + b * b); }

/* ------------------------------------------------------------------ schedule */

unsigned orc_ctf_levels(unsigned rows)
{
    /* impair.cpp:40: log2(float(rows/15)) + 1, truncated */
    return (unsigned)(log2((double)(float)(rows / 15))) + 1u;
}

void orc_num_max_iter(unsigned ctf_levels, unsigned *out)
{
    /* impair.cpp:46-52: 100 at the finest level, 15 fewer per coarser one */
    for (int i = 5; i >= 0; i--) {
        if (i >= (int)ctf_levels - 1)
            out[i] = 100;
        else
            out[i] = out[i + 1] - 15;
    }
}

void orc_gaussian_mask(float *m)
{
    /* impair.cpp:55-58: binomial [1 4 6 4 1] outer product / 256, m[i + 5j] */
    static const int b[5] = {1, 4, 6, 4, 1};
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) m[i + 5 * j] = (float)(b[i] * b[j]) / 256.f;
}

/* ------------------------------------------------------------------ lifecycle */

orc_ctx *orc_create(unsigned rows, unsigned cols, float lambda_i, float lambda_d, float mu,
                    const float *g_mask25, unsigned ctf_levels, unsigned cam_mode, float fovh,
                    float fovv, float tan_half_fovh)
{
    if (cam_mode == 0 || rows == 0 || cols == 0) return NULL;
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    if (!c) return NULL;
    c->rows = rows; c->cols = cols; c->lambda_i = lambda_i; c->lambda_d = lambda_d;
    c->mu = mu; c->ctf_levels = ctf_levels; c->cam_mode = cam_mode; c->fovh = fovh;
    c->fovv = fovv; c->tan_half = tan_half_fovh;
    memcpy(c->g_mask, g_mask25, 25 * sizeof(float));
    c->width = 640 / cam_mode;                 /* .cu:31-32 */
    c->height = 480 / cam_mode;
    /* .cu:41: pyr_levels = round(log2(width/cols)) + ctf_levels */
    c->pyr_levels = (unsigned)roundf(log2f((float)(c->width / cols))) + ctf_levels;
    if (c->pyr_levels > PDF_MAX_PYR_LEVELS) { free(c); return NULL; }
    c->nthreads = 1;
    const size_t P0 = (size_t)c->width * c->height;
    c->cap = P0 > (size_t)rows * cols ? P0 : (size_t)rows * cols;
    c->colour_wf = (float *)calloc(P0, sizeof(float));
    c->depth_wf = (float *)calloc(P0, sizeof(float));
    for (unsigned l = 0; l < c->pyr_levels; l++) {
        const size_t Pl = P0 >> (2 * l);
        for (int k = 0; k < PDF_NUM_PYR_KINDS; k++) c->pyr[k][l] = (float *)calloc(Pl, sizeof(float));
    }
    for (int i = 0; i < PDF_NUM_FIELDS; i++) c->f[i] = (float *)calloc(c->cap, sizeof(float));
    return c;
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    free(c->colour_wf);
    free(c->depth_wf);
    for (unsigned l = 0; l < PDF_MAX_PYR_LEVELS; l++)
        for (int k = 0; k < PDF_NUM_PYR_KINDS; k++) free(c->pyr[k][l]);
    for (int i = 0; i < PDF_NUM_FIELDS; i++) free(c->f[i]);
    free(c);
}

void orc_set_threads(orc_ctx *c, int n) { c->nthreads = n < 1 ? 1 : n; }

void orc_copy_new_frames(orc_ctx *c, const float *I, const float *Z)
{
    const size_t P0 = (size_t)c->width * c->height;
    memcpy(c->depth_wf, Z, P0 * sizeof(float));
    memcpy(c->colour_wf, I, P0 * sizeof(float));
    /* .cu:203-221: the previous "new" pyramid becomes "old", for all eight slots */
    for (unsigned l = 0; l < PDF_MAX_PYR_LEVELS; l++) {
        for (int k = 0; k < PDF_NUM_PYR_KINDS; k += 2) {
            float *t = c->pyr[k + 1][l];
            c->pyr[k + 1][l] = c->pyr[k][l];
            c->pyr[k][l] = t;
        }
    }
}

void orc_allocate_level(orc_ctx *c, unsigned rows_i, unsigned cols_i, unsigned level,
                        unsigned level_image)
{
    /* .cu:76-79 (the planes themselves are preallocated at full size) */
    c->level = level;
    c->level_image = level_image;
    c->rows_i = rows_i;
    c->cols_i = cols_i;
}

void orc_set_pyramid(orc_ctx *c, int kind, unsigned level, const float *in)
{
    const size_t Pl = ((size_t)c->width * c->height) >> (2 * level);
    memcpy(c->pyr[kind][level], in, Pl * sizeof(float));
}

unsigned orc_rows_i(const orc_ctx *c) { return c->rows_i; }
unsigned orc_cols_i(const orc_ctx *c) { return c->cols_i; }

void orc_get_field(const orc_ctx *c, int field, float *out)
{
    memcpy(out, c->f[field], (size_t)c->rows_i * c->cols_i * sizeof(float));
}

void orc_set_field(orc_ctx *c, int field, const float *in)
{
    memcpy(c->f[field], in, (size_t)c->rows_i * c->cols_i * sizeof(float));
}

void orc_get_pyramid(const orc_ctx *c, int kind, unsigned level, float *out)
{
    const size_t Pl = ((size_t)c->width * c->height) >> (2 * level);
    memcpy(out, c->pyr[kind][level], Pl * sizeof(float));
}

#define PIXEL_LOOP(c, N, body)                                              \
    do {                                                                    \
        const unsigned n_ = (N);                                            \
        _Pragma("omp parallel for schedule(static) num_threads(c->nthreads)") \
        for (unsigned index = 0; index < n_; index++) { body; }             \
    } while (0)

/* ------------------------------------------------------------------ K1 pyramid */

/* computePyramidLevel, .cu:331-434, for one output pixel of pyramid level `level`
 * (dims rows_i x cols_i) of the "new" pyramid */
static void pyramid_pixel(orc_ctx *c, unsigned index, unsigned level, unsigned rows_i,
                          unsigned cols_i, float inv_f_i)
{
    const float max_depth_dif = 0.1f;
    const unsigned v = index % rows_i, u = index / rows_i;
    float *col = c->pyr[PDF_PYR_COLOUR][level], *dep = c->pyr[PDF_PYR_DEPTH][level];

    if (level == 0) {                                        /* .cu:346-353 */
        col[index] = c->colour_wf[index];
        dep[index] = c->depth_wf[index];
    } else {
        const float *colp = c->pyr[PDF_PYR_COLOUR][level - 1];
        const float *depp = c->pyr[PDF_PYR_DEPTH][level - 1];
        float sumd = 0.f, sumc = 0.f, wd = 0.f, wc = 0.f;
        const unsigned centre = 2 * v + 4 * u * rows_i;     /* (2v, 2u) on the finer level */
        const float dcen = depp[centre];
        const int inner = (v > 0) && (v < rows_i - 1) && (u > 0) && (u < cols_i - 1);
        for (int k = -2; k < 3; k++)
            for (int l = -2; l < 3; l++) {
                const int vv = 2 * (int)v + k, uu = 2 * (int)u + l;
                if (!inner && !(vv >= 0 && vv < 2 * (int)rows_i && uu >= 0 && uu < 2 * (int)cols_i))
                    continue;                                 /* .cu:399 */
                const unsigned ip = (unsigned)vv + (unsigned)uu * 2 * rows_i;
                const float g = c->g_mask[12 + k + 5 * l];
                sumc += g * colp[ip];
                if (!inner) wc += g;                          /* .cu:406 */
                const float d = depp[ip];
                if ((d > 0.f) && (fabsf(d - dcen) < max_depth_dif)) { /* .cu:376-381 */
                    const float w = g * (max_depth_dif - fabsf(d - dcen));
                    wd += w;
                    sumd += w * d;
                }
            }
        dep[index] = (sumd > 0.f) ? sumd / wd : 0.f;          /* .cu:384-387, :420-423 */
        col[index] = inner ? sumc : sumc / wc;                /* interior un-normalised .cu:389 */
    }
    /* .cu:428-433 */
    const float disp_u = 0.5f * (float)(cols_i - 1);
    const float disp_v = 0.5f * (float)(rows_i - 1);
    c->pyr[PDF_PYR_XX][level][index] = ((float)u - disp_u) * dep[index] * inv_f_i;
    c->pyr[PDF_PYR_YY][level][index] = ((float)v - disp_v) * dep[index] * inv_f_i;
}

void orc_gaussian_pyramid(orc_ctx *c, unsigned levels, unsigned cam_mode)
{
    for (unsigned i = 0; i < levels; i++) {
        /* .cu:1176-1177: level dims from the camera mode */
        const unsigned cols_i = (unsigned)(640.f / ((float)cam_mode * powf(2.f, (float)i)));
        const unsigned rows_i = (unsigned)(480.f / ((float)cam_mode * powf(2.f, (float)i)));
        const float inv_f_i = 2.f * c->tan_half / (float)cols_i;
        PIXEL_LOOP(c, rows_i * cols_i, pyramid_pixel(c, index, i, rows_i, cols_i, inv_f_i));
    }
}

/* ------------------------------------------------------------------ K2-K4 */

void orc_assign_zeros(orc_ctx *c)
{
    /* .cu:439-455 */
    static const int ids[] = {PDF_DU_UPSAMP, PDF_DV_UPSAMP, PDF_DW_UPSAMP, PDF_PD_UPSAMP,
                              PDF_PUU_UPSAMP, PDF_PUV_UPSAMP, PDF_PVU_UPSAMP, PDF_PVV_UPSAMP,
                              PDF_PWU_UPSAMP, PDF_PWV_UPSAMP, PDF_DU_PREV, PDF_DV_PREV,
                              PDF_DU_NEW, PDF_DV_NEW, PDF_DW_NEW, PDF_PD, PDF_PUU, PDF_PUV,
                              PDF_PVU, PDF_PVV, PDF_PWU, PDF_PWV, PDF_DU_ACC, PDF_DV_ACC,
                              PDF_DW_ACC};
    const size_t P = (size_t)c->rows_i * c->cols_i;
    for (size_t k = 0; k < sizeof(ids) / sizeof(ids[0]); k++) memset(c->f[ids[k]], 0, P * sizeof(float));
}

static const int UPS[10] = {PDF_DU_UPSAMP, PDF_DV_UPSAMP, PDF_DW_UPSAMP, PDF_PD_UPSAMP,
                            PDF_PUU_UPSAMP, PDF_PUV_UPSAMP, PDF_PVU_UPSAMP, PDF_PVV_UPSAMP,
                            PDF_PWU_UPSAMP, PDF_PWV_UPSAMP};
static const int CARRY[10] = {PDF_DU_L, PDF_DV_L, PDF_DW_L, PDF_PD_L, PDF_PUU_L,
                              PDF_PUV_L, PDF_PVU_L, PDF_PVV_L, PDF_PWU_L, PDF_PWV_L};
/* K4 destinations, .cu:556-565 */
static const int UPDST[10] = {PDF_DU_PREV, PDF_DV_PREV, PDF_DW_NEW, PDF_PD, PDF_PUU,
                              PDF_PUV, PDF_PVU, PDF_PVV, PDF_PWU, PDF_PWV};

static void upsample_copy_pixel(orc_ctx *c, unsigned index)
{
    /* .cu:460-477: coarse (v,u) -> fine (2v, 2u); du, dv doubled */
    const unsigned rows_i = c->rows_i;
    const unsigned v = index % (rows_i / 2);
    const unsigned u = 2 * index / rows_i;
    const unsigned big = 2 * (v + u * rows_i);
    for (int k = 0; k < 10; k++) {
        const float x = c->f[CARRY[k]][index];
        c->f[UPS[k]][big] = (k < 2) ? 2.f * x : x;
    }
}

static void upsample_filter_pixel(orc_ctx *c, unsigned index)
{
    /* .cu:479-569 */
    const unsigned rows_i = c->rows_i, cols_i = c->cols_i;
    const unsigned v = index % rows_i, u = index / rows_i;
    float acc[10] = {0};
    if ((v > 1) && (v < rows_i - 2) && (u > 1) && (u < cols_i - 2)) {
        for (int k = -2; k < 3; k++)
            for (int l = -2; l < 3; l++) {
                const unsigned ii = v + k + (u + l) * rows_i;
                const float g = 4.f * c->g_mask[12 + k + 5 * l];
                for (int q = 0; q < 10; q++) acc[q] += g * c->f[UPS[q]][ii];
            }
    } else {
        float aw = 1.f;
        for (int k = -2; k < 3; k++)
            for (int l = -2; l < 3; l++) {
                const int vv = (int)v + k, uu = (int)u + l;
                const float g = 4.f * c->g_mask[12 + k + 5 * l];
                if (vv < 0 || vv >= (int)rows_i || uu < 0 || uu >= (int)cols_i) {
                    aw -= 0.25f * g;                          /* .cu:522 */
                    continue;
                }
                const unsigned ii = (unsigned)vv + (unsigned)uu * rows_i;
                for (int q = 0; q < 10; q++) acc[q] += g * c->f[UPS[q]][ii];
            }
        const float inv = 1.f / aw;                           /* .cu:542 */
        for (int q = 0; q < 10; q++) acc[q] *= inv;
    }
    for (int q = 0; q < 10; q++) c->f[UPDST[q]][index] = acc[q];
    F(PDF_DW_ACC)[index] = acc[2];                            /* .cu:568 */
}

void orc_upsample(orc_ctx *c)
{
    PIXEL_LOOP(c, c->rows_i * c->cols_i / 4, upsample_copy_pixel(c, index));
    PIXEL_LOOP(c, c->rows_i * c->cols_i, upsample_filter_pixel(c, index));
}

/* ------------------------------------------------------------------ K5 Rij */

static void rij_pixel(orc_ctx *c, unsigned index)
{
    /* .cu:725-779 */
    const unsigned rows_i = c->rows_i, cols_i = c->cols_i, L = c->level_image;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float *xo = c->pyr[PDF_PYR_XX_OLD][L], *yo = c->pyr[PDF_PYR_YY_OLD][L];
    const float *zo = c->pyr[PDF_PYR_DEPTH_OLD][L];
    const float *xn = c->pyr[PDF_PYR_XX][L], *yn = c->pyr[PDF_PYR_YY][L];
    const float *zn = c->pyr[PDF_PYR_DEPTH][L];
    float dxu = 0.f, dzu = 0.f, dxu2 = 0.f, dzu2 = 0.f;
    float dyv = 0.f, dzv = 0.f, dyv2 = 0.f, dzv2 = 0.f;
    if (u != cols_i - 1) {
        dxu = xo[index + rows_i] - xo[index];
        dzu = zo[index + rows_i] - zo[index];
        dxu2 = xn[index + rows_i] - xn[index];
        dzu2 = zn[index + rows_i] - zn[index];
    }
    if (v != rows_i - 1) {
        dyv = yo[index + 1] - yo[index];
        dzv = zo[index + 1] - zo[index];
        dyv2 = yn[index + 1] - yn[index];
        dzv2 = zn[index + 1] - zn[index];
    }
    F(PDF_RI)[index] = (fabsf(dxu) + fabsf(dzu) > 0.f) ? 2.f * rhyp(dxu, dzu) : 1.f;
    F(PDF_RJ)[index] = (fabsf(dyv) + fabsf(dzv) > 0.f) ? 2.f * rhyp(dyv, dzv) : 1.f;
    F(PDF_RI_2)[index] = (fabsf(dxu2) + fabsf(dzu2) > 0.f) ? 2.f * rhyp(dxu2, dzu2) : 1.f;
    F(PDF_RJ_2)[index] = (fabsf(dyv2) + fabsf(dzv2) > 0.f) ? 2.f * rhyp(dyv2, dzv2) : 1.f;
}

void orc_rij(orc_ctx *c) { PIXEL_LOOP(c, c->rows_i * c->cols_i, rij_pixel(c, index)); }

/* ------------------------------------------------------------------ K6 gradients */

static void imgrad_pixel(orc_ctx *c, unsigned index)
{
    /* .cu:574-627 */
    const unsigned rows_i = c->rows_i, cols_i = c->cols_i, L = c->level_image;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float *I = c->pyr[PDF_PYR_COLOUR][L], *Z = c->pyr[PDF_PYR_DEPTH][L];
    const float *ri2 = F(PDF_RI_2), *rj2 = F(PDF_RJ_2);
    const unsigned r = rows_i;
    if (u == 0) {
        F(PDF_DCU_AUX)[index] = I[index + r] - I[index];
        F(PDF_DDU_AUX)[index] = Z[index + r] - Z[index];
    } else if (u == cols_i - 1) {
        F(PDF_DCU_AUX)[index] = I[index] - I[index - r];
        F(PDF_DDU_AUX)[index] = Z[index] - Z[index - r];
    } else {
        F(PDF_DCU_AUX)[index] = (ri2[index] * (I[index + r] - I[index]) +
                                 ri2[index - r] * (I[index] - I[index - r])) /
                                (ri2[index] + ri2[index - r]);
        if (Z[index] > 0.f)
            F(PDF_DDU_AUX)[index] = (ri2[index] * (Z[index + r] - Z[index]) +
                                     ri2[index - r] * (Z[index] - Z[index - r])) /
                                    (ri2[index] + ri2[index - r]);
        else
            F(PDF_DDU_AUX)[index] = 0.f;
    }
    if (v == 0) {
        F(PDF_DCV_AUX)[index] = I[index + 1] - I[index];
        F(PDF_DDV_AUX)[index] = Z[index + 1] - Z[index];
    } else if (v == rows_i - 1) {
        F(PDF_DCV_AUX)[index] = I[index] - I[index - 1];
        F(PDF_DDV_AUX)[index] = Z[index] - Z[index - 1];
    } else {
        F(PDF_DCV_AUX)[index] = (rj2[index] * (I[index + 1] - I[index]) +
                                 rj2[index - 1] * (I[index] - I[index - 1])) /
                                (rj2[index] + rj2[index - 1]);
        if (Z[index] > 0.f)
            F(PDF_DDV_AUX)[index] = (rj2[index] * (Z[index + 1] - Z[index]) +
                                     rj2[index - 1] * (Z[index] - Z[index - 1])) /
                                    (rj2[index] + rj2[index - 1]);
        else
            F(PDF_DDV_AUX)[index] = 0.f;
    }
}

void orc_image_gradients(orc_ctx *c)
{
    PIXEL_LOOP(c, c->rows_i * c->cols_i, imgrad_pixel(c, index));
}

/* ------------------------------------------------------------------ K7 warping */

static void clamp_coords(unsigned rows_i, unsigned cols_i, float *iu, float *iv)
{
    /* .cu:656-659; a NaN coordinate is taken as 0 (undefined in the reference), exactly as
     * the HIP library does, so no value can index outside the plane */
    if (!(*iu == *iu)) *iu = 0.f;
    if (!(*iv == *iv)) *iv = 0.f;
    if (*iu < 0.f) *iu = 0.f;
    else if (*iu > (float)cols_i - 1.f) *iu = (float)cols_i - 1.f;
    if (*iv < 0.f) *iv = 0.f;
    else if (*iv > (float)rows_i - 1.f) *iv = (float)rows_i - 1.f;
}

float orc_interpolate(const float *mat, unsigned rows_i, unsigned cols_i, float iu, float iv)
{
    /* interpolatePixel, .cu:654-682 */
    clamp_coords(rows_i, cols_i, &iu, &iv);
    const unsigned su = (unsigned)ceilf(iu), fu = (unsigned)floorf(iu);
    const unsigned sv = (unsigned)ceilf(iv), fv = (unsigned)floorf(iv);
    if (su == fu && sv == fv) return mat[(long)rintf(iv + (float)rows_i * iu)];
    if (su == fu) {
        const long cu = (long)rintf(iu);
        return ((float)sv - iv) * mat[fv + rows_i * cu] + (iv - (float)fv) * mat[sv + rows_i * cu];
    }
    if (sv == fv) {
        const long cv = (long)rintf(iv);
        return ((float)su - iu) * mat[cv + rows_i * fu] + (iu - (float)fu) * mat[cv + rows_i * su];
    }
    const float top = ((float)su - iu) * mat[sv + rows_i * fu] + (iu - (float)fu) * mat[sv + rows_i * su];
    const float bot = ((float)su - iu) * mat[fv + rows_i * fu] + (iu - (float)fu) * mat[fv + rows_i * su];
    return ((float)sv - iv) * bot + (iv - (float)fv) * top;
}

float orc_interpolate_depth(const float *mat, unsigned rows_i, unsigned cols_i, float iu,
                            float iv)
{
    /* interpolatePixelDepth, .cu:684-721: any zero tap -> nearest neighbour, with
     * round-half-even on both coordinates */
    clamp_coords(rows_i, cols_i, &iu, &iv);
    const unsigned su = (unsigned)ceilf(iu), fu = (unsigned)floorf(iu);
    const unsigned sv = (unsigned)ceilf(iv), fv = (unsigned)floorf(iv);
    if (mat[sv + rows_i * su] == 0.f || mat[sv + rows_i * fu] == 0.f ||
        mat[fv + rows_i * su] == 0.f || mat[fv + rows_i * fu] == 0.f) {
        const unsigned ru = (unsigned)rintf(iu), rv = (unsigned)rintf(iv);
        return mat[rv + rows_i * ru];
    }
    return orc_interpolate(mat, rows_i, cols_i, iu, iv); /* same bilinear cases (.cu:704-719) */
}

static void warp_pixel(orc_ctx *c, unsigned index)
{
    /* .cu:629-652 */
    const unsigned rows_i = c->rows_i, cols_i = c->cols_i, L = c->level_image;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float iu = (float)u + F(PDF_DU_PREV)[index];
    const float iv = (float)v + F(PDF_DV_PREV)[index];
    const float wc = orc_interpolate(c->pyr[PDF_PYR_COLOUR][L], rows_i, cols_i, iu, iv);
    F(PDF_DCT)[index] = wc - c->pyr[PDF_PYR_COLOUR_OLD][L][index];
    F(PDF_DCU)[index] = orc_interpolate(F(PDF_DCU_AUX), rows_i, cols_i, iu, iv);
    F(PDF_DCV)[index] = orc_interpolate(F(PDF_DCV_AUX), rows_i, cols_i, iu, iv);
    const float wd = orc_interpolate_depth(c->pyr[PDF_PYR_DEPTH][L], rows_i, cols_i, iu, iv);
    F(PDF_DDT)[index] = (wd > 0.f) ? wd - c->pyr[PDF_PYR_DEPTH_OLD][L][index] : 0.f;
    F(PDF_DDU)[index] = orc_interpolate(F(PDF_DDU_AUX), rows_i, cols_i, iu, iv);
    F(PDF_DDV)[index] = orc_interpolate(F(PDF_DDV_AUX), rows_i, cols_i, iu, iv);
}

void orc_warping(orc_ctx *c) { PIXEL_LOOP(c, c->rows_i * c->cols_i, warp_pixel(c, index)); }

/* ------------------------------------------------------------------ K8 mu, steps */

static void mu_steps_pixel(orc_ctx *c, unsigned index)
{
    const unsigned rows_i = c->rows_i;
    const float ddu = F(PDF_DDU)[index], ddv = F(PDF_DDV)[index], ddt = F(PDF_DDT)[index];
    /* computeMu .cu:781-784 */
    const float mu_uv = c->mu / (1.f + 1000.f * (ddu * ddu + ddv * ddv + ddt * ddt));
    F(PDF_MU_UV)[index] = mu_uv;
    /* computeStepSizes .cu:786-808 */
    const float li = c->lambda_i, ld = c->lambda_d;
    const float ri = F(PDF_RI)[index], rj = F(PDF_RJ)[index];
    F(PDF_SIGMA_PD)[index] = 1.f / (mu_uv * (1.f + fabsf(ddu) + fabsf(ddv)) + 1e-10f);
    F(PDF_SIGMA_PUVX)[index] = 0.5f / (li * ri + 1e-10f);
    F(PDF_SIGMA_PUVY)[index] = 0.5f / (li * rj + 1e-10f);
    F(PDF_SIGMA_PWX)[index] = 0.5f / (ri * ld + 1e-10f);
    F(PDF_SIGMA_PWY)[index] = 0.5f / (rj * ld + 1e-10f);
    const unsigned v = index % rows_i, u = index / rows_i;
    float acu_r = ri + rj;
    if (u > 0) acu_r += F(PDF_RI)[index - rows_i];
    if (v > 0) acu_r += F(PDF_RJ)[index - 1];
    F(PDF_TAU_U)[index] = 1.f / (mu_uv * fabsf(ddu) + li * acu_r + 1e-10f);
    F(PDF_TAU_V)[index] = 1.f / (mu_uv * fabsf(ddv) + li * acu_r + 1e-10f);
    F(PDF_TAU_W)[index] = 1.f / (mu_uv + ld * acu_r + 1e-10f);
}

void orc_mu_and_step_sizes(orc_ctx *c)
{
    PIXEL_LOOP(c, c->rows_i * c->cols_i, mu_steps_pixel(c, index));
}

/* ------------------------------------------------------------------ K9-K12 */

static void gradient_pixel(orc_ctx *c, unsigned index)
{
    /* .cu:982-1013: forward differences of (acc + prev), zero at the last col/row */
    const unsigned rows_i = c->rows_i, cols_i = c->cols_i;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float *ua = F(PDF_DU_ACC), *va = F(PDF_DV_ACC), *wa = F(PDF_DW_ACC);
    const float *up = F(PDF_DU_PREV), *vp = F(PDF_DV_PREV);
    const float ri = F(PDF_RI)[index], rj = F(PDF_RJ)[index];
    if (u == cols_i - 1) {
        F(PDF_GRADU1)[index] = 0.f; F(PDF_GRADV1)[index] = 0.f; F(PDF_GRADW1)[index] = 0.f;
    } else {
        const unsigned n = index + rows_i;
        F(PDF_GRADU1)[index] = ri * ((ua[n] + up[n]) - (ua[index] + up[index]));
        F(PDF_GRADV1)[index] = ri * ((va[n] + vp[n]) - (va[index] + vp[index]));
        F(PDF_GRADW1)[index] = ri * (wa[n] - wa[index]);
    }
    if (v == rows_i - 1) {
        F(PDF_GRADU2)[index] = 0.f; F(PDF_GRADV2)[index] = 0.f; F(PDF_GRADW2)[index] = 0.f;
    } else {
        const unsigned n = index + 1;
        F(PDF_GRADU2)[index] = rj * ((ua[n] + up[n]) - (ua[index] + up[index]));
        F(PDF_GRADV2)[index] = rj * ((va[n] + vp[n]) - (va[index] + vp[index]));
        F(PDF_GRADW2)[index] = rj * (wa[n] - wa[index]);
    }
}

void orc_gradient(orc_ctx *c) { PIXEL_LOOP(c, c->rows_i * c->cols_i, gradient_pixel(c, index)); }

static void project_pair(float a, float b, float *oa, float *ob)
{
    /* .cu:849-859: scale onto the unit disc when outside it */
    const float m = rhyp(a, b);
    if (m < 1.f) { *oa = a * m; *ob = b * m; }
    else { *oa = a; *ob = b; }
}

static void dual_pixel(orc_ctx *c, unsigned index)
{
    /* .cu:813-887 */
    const float li = c->lambda_i, ld = c->lambda_d;
    float pd = F(PDF_PD)[index], puu = F(PDF_PUU)[index], puv = F(PDF_PUV)[index];
    float pvu = F(PDF_PVU)[index], pvv = F(PDF_PVV)[index], pwu = F(PDF_PWU)[index],
          pwv = F(PDF_PWV)[index];
    pd += F(PDF_SIGMA_PD)[index] * F(PDF_MU_UV)[index] *
          (-F(PDF_DW_ACC)[index] + F(PDF_DDT)[index] + F(PDF_DDU)[index] * F(PDF_DU_ACC)[index] +
           F(PDF_DDV)[index] * F(PDF_DV_ACC)[index]);
    puu += F(PDF_SIGMA_PUVX)[index] * li * F(PDF_GRADU1)[index];
    puv += F(PDF_SIGMA_PUVY)[index] * li * F(PDF_GRADU2)[index];
    pvu += F(PDF_SIGMA_PUVX)[index] * li * F(PDF_GRADV1)[index];
    pvv += F(PDF_SIGMA_PUVY)[index] * li * F(PDF_GRADV2)[index];
    pwu += F(PDF_SIGMA_PWX)[index] * ld * F(PDF_GRADW1)[index];
    pwv += F(PDF_SIGMA_PWY)[index] * ld * F(PDF_GRADW2)[index];
    if (fabsf(pd) > 1.f) pd = (pd > 1.f) ? 1.f : -1.f;       /* .cu:837-846 */
    F(PDF_PD)[index] = pd;
    project_pair(puu, puv, &F(PDF_PUU)[index], &F(PDF_PUV)[index]);
    project_pair(pvu, pvv, &F(PDF_PVU)[index], &F(PDF_PVV)[index]);
    project_pair(pwu, pwv, &F(PDF_PWU)[index], &F(PDF_PWV)[index]);
}

void orc_dual_variables(orc_ctx *c) { PIXEL_LOOP(c, c->rows_i * c->cols_i, dual_pixel(c, index)); }

static float div_term(const float *r, const float *p, unsigned index, unsigned stride,
                      int first, int last)
{
    /* one axis of .cu:941-979: backward differences with the reference's ends */
    if (first) return r[index] * p[index];
    if (last) return -r[index - stride] * p[index - stride];
    return r[index] * p[index] - r[index - stride] * p[index - stride];
}

static void divergence_pixel(orc_ctx *c, unsigned index)
{
    /* .cu:935-980 */
    const unsigned rows_i = c->rows_i, cols_i = c->cols_i;
    const unsigned v = index % rows_i, u = index / rows_i;
    const float *ri = F(PDF_RI), *rj = F(PDF_RJ);
    const int uf = (u == 0), ul = (u == cols_i - 1), vf = (v == 0), vl = (v == rows_i - 1);
    float du = div_term(ri, F(PDF_PUU), index, rows_i, uf, ul);
    float dv = div_term(ri, F(PDF_PVU), index, rows_i, uf, ul);
    float dw = div_term(ri, F(PDF_PWU), index, rows_i, uf, ul);
    du += div_term(rj, F(PDF_PUV), index, 1, vf, vl);
    dv += div_term(rj, F(PDF_PVV), index, 1, vf, vl);
    dw += div_term(rj, F(PDF_PWV), index, 1, vf, vl);
    F(PDF_DIVPU)[index] = du;
    F(PDF_DIVPV)[index] = dv;
    F(PDF_DIVPW)[index] = dw;
}

void orc_divergence(orc_ctx *c) { PIXEL_LOOP(c, c->rows_i * c->cols_i, divergence_pixel(c, index)); }

static void primal_pixel(orc_ctx *c, unsigned index)
{
    /* .cu:889-933 */
    const float li = c->lambda_i, ld = c->lambda_d;
    float du = F(PDF_DU_NEW)[index], dv = F(PDF_DV_NEW)[index], dw = F(PDF_DW_NEW)[index];
    const float du_old = du, dv_old = dv, dw_old = dw;
    const float tu = F(PDF_TAU_U)[index], tv = F(PDF_TAU_V)[index], tw = F(PDF_TAU_W)[index];
    const float muv = F(PDF_MU_UV)[index], pd = F(PDF_PD)[index];
    du += -tu * (muv * F(PDF_DDU)[index] * pd - li * F(PDF_DIVPU)[index]);
    dv += -tv * (muv * F(PDF_DDV)[index] * pd - li * F(PDF_DIVPV)[index]);
    dw += -tw * (-muv * pd - ld * F(PDF_DIVPW)[index]);
    /* TV-L1 intensity prox, .cu:906-923 */
    const float dct = F(PDF_DCT)[index], dcu = F(PDF_DCU)[index], dcv = F(PDF_DCV)[index];
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
    F(PDF_DU_ACC)[index] = 2.f * du - du_old;
    F(PDF_DV_ACC)[index] = 2.f * dv - dv_old;
    F(PDF_DW_ACC)[index] = 2.f * dw - dw_old;
    F(PDF_DU_NEW)[index] = du;
    F(PDF_DV_NEW)[index] = dv;
    F(PDF_DW_NEW)[index] = dw;
}

void orc_primal_variables(orc_ctx *c) { PIXEL_LOOP(c, c->rows_i * c->cols_i, primal_pixel(c, index)); }

void orc_pd_iterations(orc_ctx *c, unsigned n)
{
    for (unsigned it = 0; it < n; it++) {
        orc_gradient(c);
        orc_dual_variables(c);
        orc_divergence(c);
        orc_primal_variables(c);
    }
}

/* ------------------------------------------------------------------ K13-K15 */

static void saturate_pixel(orc_ctx *c, unsigned index)
{
    /* .cu:1018-1038 */
    float du = F(PDF_DU_NEW)[index], dv = F(PDF_DV_NEW)[index], dw = F(PDF_DW_NEW)[index];
    if (du > 1.f) du = 1.f; else if (du < -1.f) du = -1.f;
    if (dv > 1.f) dv = 1.f; else if (dv < -1.f) dv = -1.f;
    if (c->pyr[PDF_PYR_DEPTH_OLD][c->level_image][index] == 0.f) dw = 0.f;
    F(PDF_DU_NEW)[index] = du + F(PDF_DU_PREV)[index];
    F(PDF_DV_NEW)[index] = dv + F(PDF_DV_PREV)[index];
    F(PDF_DW_NEW)[index] = dw;
}

/* stable bubble sort on field (strict >), bubbleSortDev .cu:1457-1473 */
static void sort_pairs(float *fld, float *prs, int n)
{
    int go_on = 1;
    while (go_on) {
        go_on = 0;
        for (int i = 1; i < n; i++)
            if (fld[i - 1] > fld[i]) {
                float t = fld[i - 1]; fld[i - 1] = fld[i]; fld[i] = t;
                t = prs[i - 1]; prs[i - 1] = prs[i]; prs[i] = t;
                go_on = 1;
            }
    }
}

static float median_pick(const float *fld, const float *cum, float med, int n)
{
    /* .cu:1103-1110.  The scan cannot pass the end: every cumulative total is the same
     * set of positive weights summed in another order, so it exceeds med = half of u's
     * total; the bound only keeps malformed (NaN-free but adversarial) inputs in range. */
    unsigned cont = 0;
    while (cont + 1 < (unsigned)n && med > cum[cont]) cont++;
    if (cont == 0) return fld[0];
    const unsigned r = cont, l = cont - 1;
    return ((cum[r] - med) * fld[l] + (med - cum[l]) * fld[r]) / (cum[r] - cum[l]);
}

void orc_weighted_median3(int n, const float *fu, const float *fv, const float *fw,
                          const float *pres, float *out_u, float *out_v, float *out_w)
{
    float uf[9], up[9], vf[9], vp[9], wf[9], wp[9], cu[9], cv[9], cw[9];
    for (int i = 0; i < n; i++) {
        uf[i] = fu[i]; vf[i] = fv[i]; wf[i] = fw[i];
        up[i] = vp[i] = wp[i] = pres[i];
    }
    sort_pairs(uf, up, n);
    sort_pairs(vf, vp, n);
    sort_pairs(wf, wp, n);
    cu[0] = up[0]; cv[0] = vp[0]; cw[0] = wp[0];                /* .cu:1089-1095 */
    for (int i = 1; i < n; i++) {
        cu[i] = cu[i - 1] + up[i];
        cv[i] = cv[i - 1] + vp[i];
        cw[i] = cw[i - 1] + wp[i];
    }
    const float med = 0.5f * cu[n - 1];                         /* u's total, .cu:1097 */
    *out_u = median_pick(uf, cu, med, n);
    *out_v = median_pick(vf, cv, med, n);
    *out_w = median_pick(wf, cw, med, n);
}

static void filter_pixel(orc_ctx *c, unsigned index)
{
    /* .cu:1040-1148 */
    const unsigned rows_i = c->rows_i, cols_i = c->cols_i;
    const float *zo = c->pyr[PDF_PYR_DEPTH_OLD][c->level_image];
    const float d0 = zo[index];
    const unsigned v = index % rows_i, u = index / rows_i;
    if (d0 > 0.f) {
        float fu[9], fv[9], fw[9], pr[9];
        int n = 0;
        for (int k = -1; k < 2; k++)
            for (int l = -1; l < 2; l++) {
                const int r = (int)v + k, cc = (int)u + l;
                if (r < 0 || r >= (int)rows_i || cc < 0 || cc >= (int)cols_i) continue;
                const unsigned il = index + l * (int)rows_i + k;
                const float dz = d0 - zo[il], dt = F(PDF_DDT)[il];
                pr[n] = 1.f / (1.f + 5.f * (dz * dz) + 10.f * (dt * dt));
                fu[n] = F(PDF_DU_NEW)[il];
                fv[n] = F(PDF_DV_NEW)[il];
                fw[n] = F(PDF_DW_NEW)[il];
                n++;
            }
        orc_weighted_median3(n, fu, fv, fw, pr, &F(PDF_DU_L)[index], &F(PDF_DV_L)[index],
                             &F(PDF_DW_L)[index]);
    } else {
        F(PDF_DU_L)[index] = F(PDF_DU_NEW)[index];
        F(PDF_DV_L)[index] = F(PDF_DV_NEW)[index];
        F(PDF_DW_L)[index] = F(PDF_DW_NEW)[index];
    }
    F(PDF_PD_L)[index] = F(PDF_PD)[index];
    F(PDF_PUU_L)[index] = F(PDF_PUU)[index];
    F(PDF_PUV_L)[index] = F(PDF_PUV)[index];
    F(PDF_PVU_L)[index] = F(PDF_PVU)[index];
    F(PDF_PVV_L)[index] = F(PDF_PVV)[index];
    F(PDF_PWU_L)[index] = F(PDF_PWU)[index];
    F(PDF_PWV_L)[index] = F(PDF_PWV)[index];
}

void orc_filter(orc_ctx *c)
{
    PIXEL_LOOP(c, c->rows_i * c->cols_i, saturate_pixel(c, index));
    PIXEL_LOOP(c, c->rows_i * c->cols_i, filter_pixel(c, index));
}

static void motion_pixel(orc_ctx *c, unsigned index, float inv_f)
{
    /* .cu:1150-1167 */
    const unsigned L = c->level_image;
    const float z = c->pyr[PDF_PYR_DEPTH_OLD][L][index];
    if (z > 0.f) {
        const float dul = F(PDF_DU_L)[index], dvl = F(PDF_DV_L)[index], dwl = F(PDF_DW_L)[index];
        F(PDF_DX)[index] = dwl;
        F(PDF_DY)[index] = z * dul * inv_f + dwl * c->pyr[PDF_PYR_XX_OLD][L][index] / z;
        F(PDF_DZ)[index] = z * dvl * inv_f + dwl * c->pyr[PDF_PYR_YY_OLD][L][index] / z;
    } else {
        F(PDF_DX)[index] = 0.f; F(PDF_DY)[index] = 0.f; F(PDF_DZ)[index] = 0.f;
    }
}

void orc_motion_field(orc_ctx *c)
{
    /* inv_f from the finest cols at every level, .cu:1152 */
    const float inv_f = 2.f * c->tan_half / (float)c->cols;
    PIXEL_LOOP(c, c->rows_i * c->cols_i, motion_pixel(c, index, inv_f));
}

/* ------------------------------------------------------------------ driver */

void orc_solve(orc_ctx *c, const unsigned *num_max_iter)
{
    /* scene_flow_impair.cpp:95-168 */
    const unsigned lw = (unsigned)log2((double)(float)(c->width / c->cols));
    for (unsigned i = 0; i < c->ctf_levels; i++) {
        const unsigned s = (unsigned)pow(2.0, (double)(int)(c->ctf_levels - (i + 1)));
        const unsigned cols_i = c->cols / s, rows_i = c->rows / s;
        const unsigned level_image = c->ctf_levels - i + lw - 1;
        orc_allocate_level(c, rows_i, cols_i, i, level_image);
        orc_assign_zeros(c);
        if (i > 0) orc_upsample(c);
        orc_rij(c);
        orc_image_gradients(c);
        orc_warping(c);
        orc_mu_and_step_sizes(c);
        orc_pd_iterations(c, num_max_iter[i]);
        orc_filter(c);
        orc_motion_field(c);
    }
}
