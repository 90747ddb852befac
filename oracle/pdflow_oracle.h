/* CPU oracle for the PD-Flow plain solve.  TEST INFRASTRUCTURE ONLY: this code is the
 * checker that the HIP library is compared against.  Only the tests, the smoke check
 * and the CPU-baseline leg of the benchmark may load it; the product never does.
 *
 * It restates, in plain C, the per-pixel bodies of the reference's CUDA library
 * (pdflow_cudalib.cu:331-1167, :1457-1473) with the reference's kernel granularity
 * (one full pass over the level per kernel, pdflow_cudalib.cu:1272-1454) and its
 * driver sequence (scene_flow_impair.cpp:80-168).  The interface mirrors the
 * library's C-ABI stage by stage (include/pdflow_hip.h) so both can be stepped in
 * lock-step and compared plane by plane.
 */
#ifndef PDFLOW_ORACLE_H
#define PDFLOW_ORACLE_H

#include "../include/pdflow_fields.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* readParameters + allocateDevMemory (pdflow_cudalib.cu:176-193, :29-72).
 * tan_half_fovh is tan(0.5*fovh) in float, computed once by the caller so that the
 * oracle and the GPU use the same value (pdflow_cudalib.cu:428, :1152). */
orc_ctx *orc_create(unsigned rows, unsigned cols, float lambda_i, float lambda_d, float mu,
                    const float *g_mask25, unsigned ctf_levels, unsigned cam_mode,
                    float fovh, float fovv, float tan_half_fovh);
void orc_destroy(orc_ctx *c);
void orc_set_threads(orc_ctx *c, int nthreads);   /* OpenMP threads for the timed build */

void orc_copy_new_frames(orc_ctx *c, const float *I, const float *Z);   /* .cu:195-222 */
void orc_gaussian_pyramid(orc_ctx *c, unsigned levels, unsigned cam_mode); /* .cu:1172-1184 */
void orc_allocate_level(orc_ctx *c, unsigned rows_i, unsigned cols_i, unsigned level,
                        unsigned level_image);                           /* .cu:74-162 */
void orc_assign_zeros(orc_ctx *c);        /* K2 .cu:439-455 */
void orc_upsample(orc_ctx *c);            /* K3+K4 .cu:460-569 */
void orc_rij(orc_ctx *c);                 /* K5 .cu:725-779 */
void orc_image_gradients(orc_ctx *c);     /* K6 .cu:574-627 */
void orc_warping(orc_ctx *c);             /* K7 .cu:629-721 */
void orc_mu_and_step_sizes(orc_ctx *c);   /* K8 .cu:781-808 */
void orc_gradient(orc_ctx *c);            /* K9 .cu:982-1013 */
void orc_dual_variables(orc_ctx *c);      /* K10 .cu:813-887 */
void orc_divergence(orc_ctx *c);          /* K11 .cu:935-980 */
void orc_primal_variables(orc_ctx *c);    /* K12 .cu:889-933 */
void orc_pd_iterations(orc_ctx *c, unsigned n); /* n x (K9,K10,K11,K12), scene_flow_impair.cpp:138-144 */
void orc_filter(orc_ctx *c);              /* K13+K14 .cu:1018-1148 */
void orc_motion_field(orc_ctx *c);        /* K15 .cu:1150-1167 */

/* the whole coarse-to-fine solve of scene_flow_impair.cpp:95-168 with the given
 * per-level iteration counts (coarse first); the flow of the finest level is left
 * in the dx/dy/dz planes */
void orc_solve(orc_ctx *c, const unsigned *num_max_iter);

void orc_get_field(const orc_ctx *c, int field, float *out);      /* rows_i*cols_i values */
void orc_set_field(orc_ctx *c, int field, const float *in);
void orc_get_pyramid(const orc_ctx *c, int kind, unsigned level, float *out);
void orc_set_pyramid(orc_ctx *c, int kind, unsigned level, const float *in);
unsigned orc_rows_i(const orc_ctx *c);
unsigned orc_cols_i(const orc_ctx *c);

/* The schedule and geometry rules of scene_flow_impair.cpp:38-58:
 * ctf_levels, num_max_iter (coarse first), and the 5x5 binomial mask. */
unsigned orc_ctf_levels(unsigned rows);
void orc_num_max_iter(unsigned ctf_levels, unsigned *out6);
void orc_gaussian_mask(float *out25);

/* Per-pixel helpers exposed for brute-force checks. */
float orc_interpolate(const float *mat, unsigned rows_i, unsigned cols_i, float ind_u, float ind_v);
float orc_interpolate_depth(const float *mat, unsigned rows_i, unsigned cols_i, float ind_u,
                            float ind_v);
/* 3x3 weighted median of one pixel (.cu:1040-1148) on explicit tap lists: n taps of
 * (field, weight) in reference order; writes the medians of u, v, w. */
void orc_weighted_median3(int n, const float *fu, const float *fv, const float *fw,
                          const float *pres, float *out_u, float *out_v, float *out_w);

#ifdef __cplusplus
}
#endif
#endif
