// Internal interface between the C-ABI host layer (pdflow_capi.hip) and the gfx950
// kernels (pdflow_kernels.hip).  Not installed; callers use include/pdflow_hip.h.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/pdflow_fields.h"

namespace pdf {

// Pointers of one level's planes plus the pyramid planes of the level being solved.
struct LevelArgs {
    float *f[PDF_NUM_FIELDS];
    const float *I, *Z, *X, *Y;              // "new" frame pyramid at level_image
    const float *Iold, *Zold, *Xold, *Yold;  // "old" frame pyramid at level_image
    unsigned rows, cols;                     // rows_i, cols_i
    float lambda_i, lambda_d, mu;
    float inv_f;                             // 2 tan(fovh/2) / finest cols (motion field)
    float g[25];                             // Gaussian mask as given to readParameters
};

struct PyrArgs {
    const float *colp, *depp;                // finer level (level-1)
    const float *colour_wf, *depth_wf;       // ingested frame (level 0 source)
    float *col, *dep, *xx, *yy;              // output level
    unsigned rows_i, cols_i, level;
    float inv_f_i;
    float g[25];
};

// Operands of the fused primal-dual iteration.
struct PDArgs {
    const float *ri, *rj, *du_prev, *dv_prev, *mu_uv, *ddu, *ddv, *ddt, *dct, *dcu, *dcv;
    float *du, *dv, *dw, *pd;                               // updated in place
    const float *du_acc, *dv_acc, *dw_acc;                  // read (ping)
    const float *puu, *puv, *pvu, *pvv, *pwu, *pwv;
    float *du_acc_o, *dv_acc_o, *dw_acc_o;                  // written (pong)
    float *puu_o, *puv_o, *pvu_o, *pvv_o, *pwu_o, *pwv_o;
    unsigned rows, cols;
    float lambda_i, lambda_d;
};

hipError_t launch_pyramid(const PyrArgs &a, hipStream_t s);
hipError_t launch_assign_zeros(const LevelArgs &a, hipStream_t s);
hipError_t launch_upsample(const LevelArgs &a, hipStream_t s);
hipError_t launch_rij(const LevelArgs &a, hipStream_t s);
hipError_t launch_image_gradients(const LevelArgs &a, hipStream_t s);
hipError_t launch_warping(const LevelArgs &a, hipStream_t s);
hipError_t launch_mu_steps(const LevelArgs &a, hipStream_t s);
hipError_t launch_gradient(const LevelArgs &a, hipStream_t s);
hipError_t launch_dual(const LevelArgs &a, hipStream_t s);
hipError_t launch_divergence(const LevelArgs &a, hipStream_t s);
hipError_t launch_primal(const LevelArgs &a, hipStream_t s);
hipError_t launch_pd_fused(const PDArgs &a, hipStream_t s);
hipError_t launch_filter(const LevelArgs &a, hipStream_t s);
hipError_t launch_motion_field(const LevelArgs &a, hipStream_t s);

}  // namespace pdf
