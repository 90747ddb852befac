/* Plane identifiers shared by the HIP library, its C-ABI and the CPU oracle.
 *
 * Every per-level plane of the reference's CSF_cuda object (pdflow_cudalib.h:58-122)
 * gets one id, so a caller can read or overwrite any plane after any stage
 * (pdf_get_field / pdf_set_field).  All planes are FP32, column-major,
 * element (v, u) at v + u*rows_i (pdflow_cudalib.cu:343-344).
 */
#ifndef PDFLOW_FIELDS_H
#define PDFLOW_FIELDS_H

enum pdf_field_id {
    /* warped intensity / depth derivatives (pdflow_cudalib.h:76-77) */
    PDF_DCT = 0, PDF_DCU, PDF_DCV, PDF_DDT, PDF_DDU, PDF_DDV,
    /* unwarped derivatives (pdflow_cudalib.h:72-73) */
    PDF_DCU_AUX, PDF_DCV_AUX, PDF_DDU_AUX, PDF_DDV_AUX,
    /* motion-field gradients (pdflow_cudalib.h:80-82) */
    PDF_GRADU1, PDF_GRADU2, PDF_GRADV1, PDF_GRADV2, PDF_GRADW1, PDF_GRADW2,
    /* divergence of the duals (pdflow_cudalib.h:85) */
    PDF_DIVPU, PDF_DIVPV, PDF_DIVPW,
    /* step sizes (pdflow_cudalib.h:88-90) */
    PDF_SIGMA_PD, PDF_SIGMA_PUVX, PDF_SIGMA_PUVY, PDF_SIGMA_PWX, PDF_SIGMA_PWY,
    PDF_TAU_U, PDF_TAU_V, PDF_TAU_W,
    /* data-term and regulariser weights (pdflow_cudalib.h:93-95) */
    PDF_MU_UV, PDF_RI, PDF_RJ, PDF_RI_2, PDF_RJ_2,
    /* over-relaxed primal and previous-level solution (pdflow_cudalib.h:98-99) */
    PDF_DU_ACC, PDF_DV_ACC, PDF_DW_ACC, PDF_DU_PREV, PDF_DV_PREV,
    /* zero-inserted upsampling planes (pdflow_cudalib.h:102-106) */
    PDF_DU_UPSAMP, PDF_DV_UPSAMP, PDF_DW_UPSAMP, PDF_PD_UPSAMP,
    PDF_PUU_UPSAMP, PDF_PUV_UPSAMP, PDF_PVU_UPSAMP, PDF_PVV_UPSAMP,
    PDF_PWU_UPSAMP, PDF_PWV_UPSAMP,
    /* primal and dual variables (pdflow_cudalib.h:116-122) */
    PDF_DU_NEW, PDF_DV_NEW, PDF_DW_NEW, PDF_PD,
    PDF_PUU, PDF_PUV, PDF_PVU, PDF_PVV, PDF_PWU, PDF_PWV,
    /* motion field output (pdflow_cudalib.h:69) */
    PDF_DX, PDF_DY, PDF_DZ,
    /* carry planes: solution of the last solved level (pdflow_cudalib.h:109-113);
       they hold rows_i*cols_i values of the level that wrote them */
    PDF_DU_L, PDF_DV_L, PDF_DW_L, PDF_PD_L,
    PDF_PUU_L, PDF_PUV_L, PDF_PVU_L, PDF_PVV_L, PDF_PWU_L, PDF_PWV_L,
    PDF_NUM_FIELDS
};

/* Pyramid plane kinds (pdflow_cudalib.h:64-66); "old" is the earlier frame. */
enum pdf_pyr_kind {
    PDF_PYR_COLOUR = 0, PDF_PYR_COLOUR_OLD, PDF_PYR_DEPTH, PDF_PYR_DEPTH_OLD,
    PDF_PYR_XX, PDF_PYR_XX_OLD, PDF_PYR_YY, PDF_PYR_YY_OLD,
    PDF_NUM_PYR_KINDS
};

#define PDF_MAX_PYR_LEVELS 8   /* eight slots per pyramid (pdflow_cudalib.h:64-66) */
#define PDF_MAX_CTF_LEVELS 6   /* num_max_iter[6] (scene_flow_impair.h:58) */

#endif
